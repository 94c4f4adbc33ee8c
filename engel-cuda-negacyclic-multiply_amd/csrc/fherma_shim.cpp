// fherma_init / fherma_run / fherma_free over the C ABI (include/negacyclic.h).
//
// Replaces the three entry points of the reference solution
// (solve.cu:1150-1487, 1489-1528, 1530-1567) with the same signatures
// (fherma.h:45-47) and the same contract: runtime_error on invalid input or
// device failure, output shape {N, L}, state opaque and null-safe to free.
#include <string>

#include "../../include/fherma_boundary.h"
#include "../../include/negacyclic.h"

namespace {

[[noreturn]] void raise(const char* where, int status) {
    throw std::runtime_error(std::string(where) + ": " + nc_status_string(status) + ": " + nc_last_error());
}

}  // namespace

void* fherma_init(const fherma::Point& point) {
    // solve.cu:1167-1172: N a power of two, q odd with exactly L limbs
    if (point.N < 2 || point.N > 65536u || (point.N & (point.N - 1u)) != 0u)
        throw std::runtime_error("N must be a power of two from 2 to 65536");
    if (point.L == 0 || point.L > 32 || point.q.data.size() != point.L || (point.q.data[0] & 1u) == 0u)
        throw std::runtime_error("q must be an odd modulus of L limbs");
    nc_plan* plan = nullptr;
    const int rc = nc_plan_create(point.N, point.L, point.q.data.data(), &plan);
    if (rc != NC_OK) raise("fherma_init", rc);
    return plan;
}

fherma::Outputs fherma_run(void* state, const fherma::Inputs& in) {
    auto* plan = static_cast<nc_plan*>(state);
    if (!plan) throw std::runtime_error("fherma_run: null state");
    nc_plan_info info;
    int rc = nc_plan_get_info(plan, &info);
    if (rc != NC_OK) raise("fherma_run", rc);
    const size_t count = (size_t)info.N * info.L;
    if (in.a.data.size() != count || in.b.data.size() != count)
        throw std::runtime_error("a and b must contain N*L limbs");  // solve.cu:1492-1494
    fherma::Outputs out;
    out.c.shape = {(int64_t)info.N, (int64_t)info.L};
    out.c.data.resize(count);
    rc = nc_multiply(plan, in.a.data.data(), in.b.data.data(), out.c.data.data(), count);
    if (rc != NC_OK) raise("fherma_run", rc);
    return out;
}

void fherma_free(void* state) { nc_plan_destroy(static_cast<nc_plan*>(state)); }
