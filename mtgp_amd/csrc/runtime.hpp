// runtime.hpp -- internal runtime types shared by runtime.cpp and serve.cpp.
#pragma once
#include <hip/hip_runtime_api.h>

#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "mtgp32-fast.h"
#include "mtgp64-fast.h"
#include "mtgp_hip.h"

#include "common.hpp"
#include "gf2x.hpp"

namespace mtgp {

struct Error : std::runtime_error {
    explicit Error(const std::string &m);
};
void hip_check(hipError_t e, const char *what);

// Per-device state: the parameter table in device memory and launch
// limits from the occupancy calculator.
struct Context {
    int device = 0;
    int cus = 0;
    size_t hbm_bytes = 0;
    std::string arch;
    DevParam *d_params = nullptr;
    int n_params = 0, param_capacity = 0;
    // Occupancy limit [bits==64][class][layout] and the default layout.
    int blocks_per_cu[2][3][2] = {};
    int jump_blocks_per_cu[2][3] = {};
    int layout[2][3] = {{1, 1, 1}, {1, 1, 1}};
    // Resident workgroups per CU used by default (measured best for the
    // generation kernels: 16 of the 32 the occupancy limit allows).
    int default_bpc = 16;
    // Automatic segmentation cuts the work of a plan into about this many
    // segments per resident workgroup slot: more rounds balance the tail
    // better, each extra segment costs one jump-ahead (1.5 measured best for
    // uneven MTGP32-11213 batches at 16 workgroups per CU).
    double segment_rounds = 1.5;
    struct Custom {
        DevParam rec;
        unsigned char sha1[20];
    };
    std::vector<Custom> customs;
    std::mutex mu;

    std::map<std::pair<const void *, int>, size_t> reserve_cache;
    // Dynamic LDS that limits `fn` to `per_cu` resident workgroups per CU
    // (0 when the kernel's own limit is already at most per_cu).
    size_t lds_reserve(const void *fn, int per_cu);
    int slot_of(const DevParam &rec, const unsigned char *sha1);
    const DevParam &param(int slot) const;
    void sha1_of(int slot, unsigned char out[20]) const;
};

Context &context();  // current device's context (throws without a GPU)
void select_device(int dev);

DevParam dev_param_of(const mtgp32_params_fast_t &p);
DevParam dev_param_of(const mtgp64_params_fast_t &p);

// Characteristic polynomial (as a modulus) of a parameter set, computed by
// Berlekamp-Massey and checked against the set's poly_sha1; cached.
std::shared_ptr<gf2::Modulus> modulus_for(const DevParam &d, const unsigned char *sha1);
std::shared_ptr<gf2::Modulus> modulus_for(Context &c, int slot);

// Per-residue coefficient masks of up to kJumpBatch jump polynomials
// (layout in common.hpp), appended to `out`; returns the chunk count.
int append_jump_masks(const std::vector<const std::vector<uint32_t> *> &polys, std::vector<uint64_t> &out);

// A generator as the planner sees it.
struct GenRef {
    int bits = 32, mexp = 0, n = 0, slot = 0;
    void *state_in = nullptr;   // current window (device)
    void *state_out = nullptr;  // where the window after the plan goes
};

uint64_t auto_segment(const Context &c, uint64_t total_numbers, int n_requests, int slots);

// A list of device Items plus the launches that run them.
struct Plan {
    struct Launch {
        int bits, cls, layout, first, count, grid;
        size_t dyn_lds;
    };
    Context *ctx = nullptr;
    // Start states of segments: 0 = jump inside the generation kernel
    // (Item.poly), 1 = a batched jump launch first (JumpItem) writing them
    // to a state arena, so the generation launch is jump-free.
    int jump_mode = 1;
    std::vector<Item> items;
    std::vector<Launch> launches;
    std::vector<JumpItem> jitems;
    std::vector<Launch> jlaunches;  // layout unused
    JumpItem *d_jitems = nullptr;
    uint64_t *d_masks = nullptr;
    char *d_jstates = nullptr;
    std::vector<int> commits;  // generators whose state moves
    Item *d_items = nullptr;
    uint32_t *d_polys = nullptr;
    uint32_t *d_queue = nullptr;
    uint64_t segment_used = 0;
    double plan_ms = 0;

    void build(Context &c, const std::vector<GenRef> &gens, const std::vector<mtgp_request_t> &reqs,
               uint64_t segment);
    void launch(hipStream_t s);
    void launch_part(int part, hipStream_t s);  // 0 jumps, 1 generation
    // Overrides the step layout and/or the resident workgroups per CU of
    // every launch (-1 keeps the current value).
    void configure(int layout, int blocks_per_cu);
    void free_device();
    Plan() = default;
    Plan(const Plan &) = delete;
    Plan &operator=(const Plan &) = delete;
    ~Plan();
};

}  // namespace mtgp
