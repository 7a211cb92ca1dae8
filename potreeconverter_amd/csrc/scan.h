// Device-wide exclusive prefix sum (decoupled look-back).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcg {
size_t scan_workspace_bytes(int64_t n);
// In-place exclusive scan of n uint32 values; *d_total (optional) receives the sum.
int exclusive_scan_u32(uint32_t* d_data, int64_t n, void* d_workspace, uint64_t* d_total, hipStream_t stream);
}  // namespace pcg
