// kernels.hpp -- launch interface of kernels.hip.
#pragma once
#include <hip/hip_runtime_api.h>
#include "common.hpp"

#include <string>

namespace mtgp {

// Kernel class for a state of n words (-1 if none fits), assuming
// N - pos >= 256 (every table set).
int kernel_class(int bits, int n);
// The same for any set: 64-bit states with N <= 176 and 128 <= N - pos <
// 256 (the MTGP64 dynamic-creator sets) get class 3, whose kernels step 128
// words at a time (lane-strided layout only, batched start-state jumps).
int kernel_class(int bits, int n, int pos);
constexpr int kClasses = 4;  // classes 0..2 of both word sizes, 3 for 64-bit
// Step layouts: 0 = each lane computes 4 consecutive words and stores them
// with one 128-bit store; 1 = lane l computes words l + 64i (conflict-free
// LDS, 256-byte coalesced dword stores); 2 = layout 1 with two waves per
// cell, thread l computing words l + 128i of 512-word steps (256-word steps
// of two words per lane for class 0); 3 = four waves per cell, 1024-word
// steps (class 2; 512-word steps of two words per lane for class 1).  Layouts 2 and 3 need N - pos of at least the step and
// batched start states.
constexpr int kLayouts = 4;
// Whether (bits, cls) has a kernel of that layout, and its words per step.
bool layout_supported(int bits, int cls, int layout);
int layout_step_words(int bits, int cls, int layout);
// Waves per workgroup (per cell) of a layout.
inline int layout_waves(int layout) { return layout == 3 ? 4 : layout == 2 ? 2 : 1; }
// Kernel entry (for occupancy queries) and its workgroup size.
const void *kernel_fn(int bits, int cls, int layout);
int kernel_threads(int bits, int cls, int layout);
// Its template name, as profilers print it (e.g. mtgp32_items_kernel<351,3,1,1,4>).
std::string kernel_name(int bits, int cls, int layout);
// Runs items[0, n_items) (device array, all of word size `bits` and class
// <= cls) with `grid` persistent workgroups.  `queue` is a device pair of
// uint32_t counters that is zero on entry and zero again on exit.
// `dyn_lds` bytes of unused dynamic LDS per workgroup cap how many
// workgroups a CU holds at once (even placement of one-item-per-wave grids).
hipError_t launch_items(int bits, int cls, int layout, const Item *items, int n_items, const DevParam *params,
                        uint32_t *queue, int grid, size_t dyn_lds, hipStream_t s);

// Batched jump kernel (JumpItem, common.hpp).
const void *jump_kernel_fn(int bits, int cls);
hipError_t launch_jumps(int bits, int cls, const JumpItem *items, int n_items, const DevParam *params,
                        uint32_t *queue, int grid, size_t dyn_lds, hipStream_t s);

// Matrix-core jumps (MJumpGroup, common.hpp), MTGP32 sets with n <= 1391
// and n - pos >= 256: first the bit planes of every group (one workgroup per
// group), then the products (one workgroup per MJumpItem; every item of a
// launch has jumps in exactly row_tiles of its 32-jump row tiles, 1..2).
hipError_t launch_planes32(const MJumpGroup *groups, int n_groups, const DevParam *params, hipStream_t s);
hipError_t launch_jump_mfma32(const MJumpGroup *groups, const MJumpItem *items, int n_items, int row_tiles,
                              hipStream_t s);

// init_kernels.hip: one wave per InitItem (n <= 1727 words).
hipError_t launch_init(const InitItem *items, int n_items, hipStream_t s);

// Writes `value` to n_words 32-bit words at dst with about `waves`
// single-wave workgroups, each storing its own contiguous rows.
hipError_t launch_fill(void *dst, uint64_t n_words, uint32_t value, int waves, hipStream_t s);
// Copies n16 16-byte units src -> dst (device memory, 16-byte aligned).
hipError_t launch_copy16(const void *src, void *dst, uint64_t n16, int grid, hipStream_t s);
// Combines split windows: for each JumpItem (one workgroup each), the
// window of state_bytes at states_out becomes the XOR of the in_parts
// windows at state_in, in_part_bytes apart (JumpItem.masks unused).
hipError_t launch_combine(const JumpItem *items, int n_items, hipStream_t s);

}  // namespace mtgp
