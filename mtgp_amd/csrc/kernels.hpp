// kernels.hpp -- launch interface of kernels.hip.
#pragma once
#include <hip/hip_runtime_api.h>
#include "common.hpp"

namespace mtgp {

// Kernel class for a state of n words (-1 if none fits).
int kernel_class(int bits, int n);
// Static LDS bytes per workgroup of a class (one wave per workgroup).
int kernel_lds_bytes(int bits, int cls);
// Kernel entry (for occupancy queries).
const void *kernel_fn(int bits, int cls);
// Runs items[0, n_items) (device array, all of word size `bits` and class
// <= cls) with `grid` persistent workgroups.  `queue` is a device pair of
// uint32_t counters that is zero on entry and zero again on exit.
hipError_t launch_items(int bits, int cls, const Item *items, int n_items, const DevParam *params,
                        uint32_t *queue, int grid, hipStream_t s);

}  // namespace mtgp
