// runtime.hpp -- internal runtime types shared by runtime.cpp and serve.cpp.
#pragma once
#include <hip/hip_runtime_api.h>

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
    int blocks_per_cu[2][3] = {{1, 1, 1}, {1, 1, 1}};  // [bits==64][class]
    struct Custom {
        DevParam rec;
        unsigned char sha1[20];
    };
    std::vector<Custom> customs;
    std::mutex mu;

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

// A generator as the planner sees it.
struct GenRef {
    int bits = 32, mexp = 0, n = 0, slot = 0;
    void *state_in = nullptr;   // current window (device)
    void *state_out = nullptr;  // where the window after the plan goes
};

uint64_t auto_segment(const Context &c, uint64_t total_numbers);

// A list of device Items plus the launches that run them.
struct Plan {
    struct Launch {
        int bits, cls, first, count, grid;
    };
    Context *ctx = nullptr;
    std::vector<Item> items;
    std::vector<Launch> launches;
    std::vector<int> commits;  // generators whose state moves
    Item *d_items = nullptr;
    uint32_t *d_polys = nullptr;
    uint32_t *d_queue = nullptr;
    uint64_t segment_used = 0;
    double plan_ms = 0;

    void build(Context &c, const std::vector<GenRef> &gens, const std::vector<mtgp_request_t> &reqs,
               uint64_t segment);
    void launch(hipStream_t s);
    void free_device();
    Plan() = default;
    Plan(const Plan &) = delete;
    Plan &operator=(const Plan &) = delete;
    ~Plan();
};

}  // namespace mtgp
