// Library-internal entry points shared by the C ABI, the multi-N batch
// graph and the server (not exported).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "nc_plan.hpp"

namespace nc {

// Records `msg` as this thread's last error and returns `status`.
int set_error(int status, const std::string& msg);

// Configuration a call with `batch` products uses (tuned, forced or default).
Choice choose_for(const nc_plan* p, uint32_t batch);

// Device workspace bytes one product needs per operand (residue rows).
size_t workspace_bytes_per_product(const nc_plan* p);

// The product pipeline on caller-provided workspace: ws_a and ws_b each hold
// `batch` products' residue rows (workspace_bytes_per_product each).
// Enqueues only: no allocation, no synchronisation (safe under graph capture).
int run_pipeline_ws(nc_plan* p, void* ws_a, void* ws_b, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c,
                    uint32_t batch, hipStream_t s, const Choice& c);

// Validates `batch` for this plan (count and N limits) without touching
// the plan's workspace.
int validate_batch(const nc_plan* p, uint32_t batch);


}  // namespace nc
