// Entry point of the kernel emulation library (tests/emu): runs a list of
// Items through the library's own kernels.hip compiled as host C++.
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

extern "C" int emu_run_items(int bits, int cls, int layout, const mtgp::Item *items, int n_items, int grid) {
    static std::vector<mtgp::DevParam> params;
    if (params.empty())
        for (int k = 0; k < mtgp::kNumSets; k++)
            params.push_back(mtgp::host_param(k));
    uint32_t queue[2] = {0, 0};
    if (mtgp::launch_items(bits, cls, layout, items, n_items, params.data(), queue, grid, 0, nullptr) != hipSuccess)
        return -1;
    return (queue[0] == 0 && queue[1] == 0) ? 0 : -2;  // the counters must reset themselves
}

extern "C" int emu_run_jumps(int bits, int cls, const mtgp::JumpItem *items, int n_items, int grid) {
    static std::vector<mtgp::DevParam> params;
    if (params.empty())
        for (int k = 0; k < mtgp::kNumSets; k++)
            params.push_back(mtgp::host_param(k));
    uint32_t queue[2] = {0, 0};
    if (mtgp::launch_jumps(bits, cls, items, n_items, params.data(), queue, grid, 0, nullptr) != hipSuccess)
        return -1;
    return (queue[0] == 0 && queue[1] == 0) ? 0 : -2;
}

extern "C" int emu_run_init(const mtgp::InitItem *items, int n_items) {
    return mtgp::launch_init(items, n_items, nullptr) == hipSuccess ? 0 : -1;
}

// The same two with a caller-supplied parameter table (sets outside the
// library's tables, e.g. dynamic-creator sets).
extern "C" int emu_run_items_p(int bits, int cls, int layout, const mtgp::Item *items, int n_items, int grid,
                               const mtgp::DevParam *params) {
    uint32_t queue[2] = {0, 0};
    if (mtgp::launch_items(bits, cls, layout, items, n_items, params, queue, grid, 0, nullptr) != hipSuccess)
        return -1;
    return (queue[0] == 0 && queue[1] == 0) ? 0 : -2;
}

extern "C" int emu_run_jumps_p(int bits, int cls, const mtgp::JumpItem *items, int n_items, int grid,
                               const mtgp::DevParam *params) {
    uint32_t queue[2] = {0, 0};
    if (mtgp::launch_jumps(bits, cls, items, n_items, params, queue, grid, 0, nullptr) != hipSuccess)
        return -1;
    return (queue[0] == 0 && queue[1] == 0) ? 0 : -2;
}

// Matrix-core jumps: the planes kernel over every group, then the products
// over the items (all with row_tiles row tiles).
extern "C" int emu_run_mjumps(const mtgp::MJumpGroup *groups, int n_groups, const mtgp::MJumpItem *items, int n_items,
                              int row_tiles) {
    static std::vector<mtgp::DevParam> params;
    if (params.empty())
        for (int k = 0; k < mtgp::kNumSets; k++)
            params.push_back(mtgp::host_param(k));
    if (mtgp::launch_planes32(groups, n_groups, params.data(), nullptr) != hipSuccess)
        return -1;
    if (mtgp::launch_jump_mfma32(groups, items, n_items, row_tiles, nullptr) != hipSuccess)
        return -2;
    return 0;
}
