// runtime.hpp -- internal runtime types shared by runtime.cpp and serve.cpp.
#pragma once
#include <hip/hip_runtime_api.h>

#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <stdexcept>
#include <string>
#include <tuple>
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
    // Occupancy limit [bits==64][class][layout] and the default layout
    // (class 3 exists for 64-bit only; kernels.hpp).
    int blocks_per_cu[2][4][4] = {};  // [bits == 64][class][layout]
    int jump_blocks_per_cu[2][4] = {};
    // Multi-wave cells by default where they measured faster: on 64-set
    // batches of MTGP32-23209 / MTGP64-23209, two-wave cells (layout 2) 3.32
    // / 4.09 TB/s per step against 2.60 / 3.86-3.95 with single-wave cells;
    // MTGP32-44497 / MTGP64-44497, four-wave cells (layout 3) on 768 / 512
    // cells 3.03 / 3.76 against 2.22 / 3.49 with two-wave cells and 1.78 /
    // 2.83 with single-wave ones (more waves per CU from fewer cells, so
    // fewer start-state jumps).  Batched MTGP32-11213 keeps
    // single-wave cells: two-wave ones measured 4.16 / 4.25 / 4.04 TB/s at 6
    // / 8 / 12 per CU against 4.21.
    int layout[2][4] = {{1, 2, 3, 1}, {2, 3, 2, 1}};
    // The same for the step-time model's one-round grids, where MTGP32-11213
    // also gains from two-wave cells (one stream of 2^24 + 2^24 numbers:
    // 0.500 against 0.467 TB/s per step, generation 0.074 against 0.093 ms)
    // while several rounds of them lose (4096 substreams: 4.21 against 4.46).
    int layout_one_round[2][4] = {{2, 2, 3, 1}, {2, 3, 2, 1}};
    // Resident workgroups per CU used by default.  The generation kernel is
    // held back by the board's power budget (its clock drops to 1.6-2.0 GHz
    // under full store traffic), so more waves do not generate faster but
    // do mean more cells, hence more jumps: on the batched MTGP32-11213
    // workload 12 per CU measured 4.49 ms per step (generation 3.90 ms)
    // against 4.60-4.63 ms at 16 and 4.71 ms at 8.
    int default_bpc = 12;
    // Resident workgroups per CU for generation launches per [bits==64][class]
    // (until the autotuner measures better; MTGP_BLOCKS_PER_CU sets them
    // all).  default_bpc everywhere but the MTGP64 128-word-step class
    // (dynamic-creator sets, N = 176): 8, which halves the cells' jumps
    // there; its 64-set batch measured 4207-4212 GB/s per step at 8
    // against 4032-4041 at 12 and 3557-3574 at 16, while the other classes
    // are even or slower at 8.
    int bpc[2][4] = {{12, 12, 12, 12}, {12, 12, 12, 8}};
    // Automatic segmentation cuts the work of a plan into about this many
    // segments per resident workgroup slot: more rounds balance the tail
    // better, each extra segment costs one jump-ahead (1.5 measured best for
    // uneven MTGP32-11213 batches at 16 workgroups per CU; used when the
    // two-size grid is off).
    double segment_rounds = 1.5;
    // Two-size automatic cells: the first `segment_split` of each
    // generator's positions in cells sized for one round over all slots, the
    // rest in cells sized for a second round (so the last round's cells are
    // small and fill the tail); 0 = equal cells, segment_rounds rounds.
    // 0.76 measured best on the batched MTGP32-11213 workload at 12
    // workgroups per CU with device-scope output stores (jump + generation
    // 4.19 ms; 0.75: 4.19-4.23, 0.77: 4.21-4.23, 0.74 and 0.78: 4.26-4.30,
    // 0.8: 4.30-4.35, 0.7: 4.32-4.34, 0.85: 4.47-4.52); MTGP64-DC batches
    // gain 1.5% over 0.8 and the other batches are even.  (With wave-scope
    // stores 0.8 had been best.)
    double segment_split = 0.76;
    // Small cells of the two-size grid per resident slot (their count is
    // about small_rounds * slots; each costs a jump).
    double small_rounds = 1.0;
    // Jump-cost-aware grid.  With rho = (jump cost) / (cell size of one
    // round over every slot) <= two_size_rho the two-size grid above is
    // used; beyond it, one round of equal cells, w per CU, with w chosen to
    // minimise the modelled step time (w a multiple of 4, the SIMDs of a CU)
    //   jump_floor_ms + (w * CUs - generators) * t_jump
    //     + bytes / (rate * (1 - exp(-w / w0)))
    // with the class's measured jump time, saturated generation rate and
    // wave scale (grid_model_of; MTGP32 waves write half the bytes per step
    // of MTGP64 ones and need more company).  It picks the measured best of
    // 4, 8 and 12 cells per CU for every 64-set MTGP32/MTGP64 batch (e.g.
    // MTGP32-23209: 1.66 TB/s at 4, 2.44 at 8, 2.52 at 12) and 4 for one
    // MTGP32-11213 stream.
    double two_size_rho = 0.15;
    double jump_floor_ms = 0.15;
    // Two-phase launches of two-size grids (MTGP_OVERLAP=1).  Off by
    // default: on the batched MTGP32-11213 workload the side stream's jumps
    // share the generation kernel's power budget rather than hide behind its
    // stores (4.8 ms per step against 4.6 ms serial).
    bool overlap = false;
    // Batched jumps are split into partial jumps (chunk ranges, XORed when
    // the generation kernel loads the start window) until a jump launch has
    // about this many waves per CU: a jump launch is one round of waves, so
    // its duration is one wave's share of the chunks.
    int jump_waves_per_cu = 16;
    // Most partial jumps one batched jump is split into (each reads and
    // XORs its partial window when the generation kernel loads the cell).
    // 16 measured best: one MTGP32-11213 stream (1017 jumps) 0.195 ms per
    // jump launch against 0.25 ms at 8; batches unchanged.
    int jump_max_parts = 16;
    // A generator's jumps of one plan run on the matrix cores (MJumpGroup)
    // when there are at least this many and its set is an MTGP32 one;
    // 0 = never.  Measured on MI355X: one MTGP32-11213 stream of 2^24 + 2^24
    // numbers (1016 jumps) 0.115 ms per jump launch against 0.19 ms; the
    // 128-generator batch (45 jumps each) 0.53 against 0.51 ms, and the
    // generation launch after a matrix-core jump launch runs 0.2 ms slower
    // (the fp4 products draw the power the store-bound generation runs
    // short of), so batches keep the batched jump kernel.
    int jump_mfma_min = 256;
    struct Custom {
        DevParam rec;
        unsigned char sha1[20];
    };
    // Reserved to param_capacity - kNumSets when the context is made, so
    // param() references stay valid while slot_of appends under `mu`.
    std::vector<Custom> customs;
    std::mutex mu;
    // Guards the planner defaults above (layout, layout_one_round, bpc and
    // the tuning knobs): plan builds read them under a shared lock,
    // mtgp_tuning_set and the autotuner write them under an exclusive one.
    mutable std::shared_mutex tune_mu;

    // mtgp32_fill / mtgp64_fill reuse one plan and grow-only device
    // buffers (start and end windows, an output staging buffer for host
    // destinations), so repeated calls allocate nothing; calls on one
    // device serialise on fill_mu.  Kept until the process ends.
    std::mutex fill_mu;
    struct Plan *fill_plan = nullptr;
    char *fill_states = nullptr, *fill_out = nullptr;
    size_t fill_states_cap = 0, fill_out_cap = 0;
    // Pinned staging of the windows and of host outputs up to 64 MiB (grows
    // to the largest such call and is kept until the process ends).
    char *fill_h = nullptr;
    size_t fill_hcap = 0;
    // The shape fill_plan was built for.  A call of the same shape launches
    // it again without planning: its cells, jumps and output addresses do
    // not depend on the start window, which the planes and jump kernels
    // read when they run.
    struct FillShape {
        int slot = -1, bits = 0, kind = -1;
        size_t n = 0;
        const void *out = nullptr;
        const char *states = nullptr;
        uint64_t epoch = ~0ull;
        bool operator==(const FillShape &o) const {
            return slot == o.slot && bits == o.bits && kind == o.kind && n == o.n && out == o.out &&
                   states == o.states && epoch == o.epoch;
        }
    };
    FillShape fill_shape;
    // The single-jump path (mtgp*_fast_jump, mtgp_genset_jump) keeps one
    // grow-only device workspace and its pinned staging copy; calls on one
    // device serialise on jump_mu.  Kept until the process ends.
    std::mutex jump_mu;
    char *jump_d = nullptr, *jump_h = nullptr;
    size_t jump_dcap = 0, jump_hcap = 0;
    // Bumped under an exclusive tune_mu whenever a planner default may
    // change, so that a plan kept for reuse knows it is stale.
    uint64_t tune_epoch = 0;

    std::map<std::pair<const void *, int>, size_t> reserve_cache;
    // Dynamic LDS that limits `fn` to `per_cu` resident workgroups per CU
    // (0 when the kernel's own limit is already at most per_cu).
    size_t lds_reserve(const void *fn, int threads, int per_cu);
    int slot_of(const DevParam &rec, const unsigned char *sha1);
    const DevParam &param(int slot) const;
    // Kernel class of a parameter slot (from its N and pos).
    int class_of(int slot) const;
    // Layouts the slot's set admits, as a bit mask (bit l: layout l); the
    // multi-wave layouts need N - pos of at least their step.
    int layouts_ok(int slot) const;
    // `lay` where the mask allows it, else the next narrower layout (3 -> 2
    // -> 1).
    static int layout_within(int lay, int mask);
    // The default layout of (bits, class) within the mask, for a one-round
    // grid or any other.
    int layout_for(int bits, int cls, int mask, bool one_round = false) const;
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

// Per-residue 16-bit form of one jump polynomial given as 32-bit
// coefficient words: m[c * 64 + rho] bit q = coefficient of
// x^(c * kJumpChunk + 64 q + rho).  And back.
std::vector<uint16_t> jump_mask16(const std::vector<uint32_t> &words);
std::vector<uint32_t> words_from_mask16(const std::vector<uint16_t> &m, int nwords);
// Per-residue coefficient masks of up to kJumpBatch jumps (layout in
// common.hpp), appended to `out`; returns the chunk count.
int append_jump_masks16(const std::vector<const std::vector<uint16_t> *> &ms, std::vector<uint64_t> &out);
int append_jump_masks(const std::vector<const std::vector<uint32_t> *> &polys, std::vector<uint64_t> &out);

// A generator as the planner sees it.
struct GenRef {
    int bits = 32, mexp = 0, n = 0, slot = 0;
    void *state_in = nullptr;   // current window (device)
    void *state_out = nullptr;  // where the window after the plan goes
};

uint64_t auto_segment(const Context &c, uint64_t total_numbers, int n_requests, int slots);
// Cost of one batched jump-ahead, in numbers the same kernel class
// generates in the same chip time: a jump is about MEXP * N / 2 word XORs
// (on two planes for MTGP64, whose numbers also take twice as long to
// generate).  Measured on MI355X at about 0.026 * MEXP * N numbers for
// MTGP32-11213 and MTGP64-23209 alike.
double jump_cost_numbers(int mexp, int n);
// The grid model's inputs (Context::two_size_rho): for one kernel class,
// or numbers-weighted means over the generators of a plan.
struct GridModel {
    double bytes = 0;     // output bytes of the plan
    double jump_ns = 0;   // chip time of one batched jump
    double rate = 4700;   // saturated generation rate, GB/s
    double w0 = 5.0;      // waves per CU at which the rate is 63% of it
    double cell_waves = 1;        // waves per cell on the one-round grid
    double cell_waves_multi = 1;  // the same on multi-round grids
};
// Measured on MI355X per (word size, state class); classes without a
// measurement use jump_ns = 4.5e-5 * MEXP * N * (bits / 32) / 2 (the jump's
// word XORs at the rate measured for MTGP32-11213 and MTGP64-23209).
GridModel grid_model_of(int bits, int cls, int mexp, int n);
// Cell sizes for automatic segmentation: S1 for the first `frac` of each
// generator's positions, S2 after (frac = 0: one size).  `jump_cost` is the
// numbers-weighted mean jump_cost_numbers of the plan's generators: when a
// jump costs much less than a cell of one round over every slot, the
// two-size grid (two rounds); otherwise one size, with fewer cells the
// dearer a jump is.  `slots` counts resident workgroups; the multi-round
// grids size their rounds in waves (slots / model.cell_waves cells), so
// multi-wave cells keep the waves per CU of single-wave ones.  Returns true
// for the one-round grid of the step-time model.
bool auto_grid(const Context &c, uint64_t total_numbers, int n_requests, int n_gens, int slots, double jump_cost,
               const GridModel &model, uint64_t &S1, uint64_t &S2, double &frac);

// Jump polynomials x^pos mod P kept per parameter slot for positions on a
// grid of `cell` numbers (pos = k * cell): table T1[k] = x^(k cell) for
// k < B = kmax + 1, and T2[j] = x^(j B cell) for j < 64, so any k < 64 B
// costs at most one multiplication; x^(2^i cell) composes larger k.
// Read-only after construction, so planners on several threads may share
// it.
class JumpCache {
  public:
    // far = false leaves out the tables for positions beyond kmax cells
    // (lookups there then fail and the planner computes those cells).
    JumpCache(Context &c, const std::vector<int> &slots, uint64_t cell, int kmax, bool far = true);
    // One cache per (device, cell, kmax, slots) for the whole process.
    static std::shared_ptr<JumpCache> shared(Context &c, const std::vector<int> &slots, uint64_t cell, int kmax,
                                             bool far = true);
    uint64_t cell() const { return cell_; }
    // x^pos mod P_slot in the per-residue mask form (jump_mask16), when pos
    // is a positive multiple of the cell size and the slot is cached; false
    // otherwise.  `cheap` limits the lookup to table entries (no product).
    bool get_masks(int slot, uint64_t pos, std::vector<uint16_t> &m16, bool cheap = false) const;
    // The same as coefficient words.
    bool get(int slot, uint64_t pos, std::vector<uint32_t> &words, bool cheap = false) const;
    // The table entry x^pos mod P_slot as coefficient words, if pos is one
    // (a positive multiple of the cell size below the first table's end);
    // nullptr otherwise.
    const std::vector<uint32_t> *table_words(int slot, uint64_t pos) const;

  private:
    struct Entry {
        std::shared_ptr<gf2::Modulus> mod;
        int words32 = 0;
        std::vector<std::vector<uint16_t>> t1;  // [k] = x^(k cell) in mask form, k < B
        std::vector<std::vector<uint32_t>> t1w; // the same as coefficient words
        std::vector<gf2::Poly> t2p;              // [j] = x^(j B cell)
        std::vector<gf2::Poly> pow2;             // [i] = x^(2^i cell)
    };
    uint64_t cell_;
    uint64_t B_ = 0;
    std::map<int, Entry> entries_;
};

struct PlanOptions {
    uint64_t segment = 0;          // numbers per grid cell (0: automatic)
    uint64_t discard_max = 65536;  // gaps up to this size are stepped over, not jumped
    const JumpCache *cache = nullptr;
    hipStream_t stream = nullptr;  // upload stream
    bool sync = true;              // wait for the upload before returning
    int jump_mode = -1;            // 0 fused, 1 batched, -1 the plan's default
    int max_parts = 0;             // waves a batched jump is split over, at most (0: jump_max_parts)
    int overlap = -1;              // two-phase launch (see Plan): 1 on, 0 off, -1 default
};

// A list of device Items plus the launches that run them.
//
// Each generator's requests occupy consecutive positions of its stream
// (skips leave gaps).  The positions are cut on a grid of S numbers into
// cells; a cell is one Item: a jump to the cell's start (or the generator's
// state for the first cell) and then its segments in order -- pieces of
// requests, or numbers stepped over -- so consecutive requests need no jump
// between them.
//
// Two-phase launch (overlap): with the two-size grid, the large cells (one
// per resident workgroup, the first round) form phase 0 and the small cells
// phase 1.  launch() runs phase 0's jumps, then phase 0's generation on the
// caller's stream while phase 1's jumps, then phase 1's generation, run on
// a side stream; the caller's stream waits for the side stream at the end.
// Phase 1's jumps then run beside generation and phase 1's generation fills
// phase 0's tail; measured on MI355X the jumps share the generation
// kernel's power budget instead of hiding behind it, so the order is off by
// default (set_overlap).  launch_part runs the merged launches (one jump
// and one generation launch per class over both phases: the serial order).
struct Plan {
    struct Launch {
        int bits, cls, layout, first, count, grid;
        size_t dyn_lds;
        int phase;
        int layouts_ok;  // bit mask of the layouts allowed for these items
    };
    Context *ctx = nullptr;
    // Start states of cells: 0 = jump inside the generation kernel
    // (Item.poly), 1 = a batched jump launch first (JumpItem) writing them
    // to a state arena, so the generation launch is jump-free.
    int jump_mode = 1;
    std::vector<Item> items;
    std::vector<Segment> segs;
    std::vector<Launch> launches;
    std::vector<JumpItem> jitems;
    std::vector<Launch> jlaunches;  // layout unused
    // The same work merged over the phases: one range per (word size,
    // class), run by launch_part (serial order).
    std::vector<Launch> mlaunches, mjlaunches;
    std::vector<int> commits;       // generators whose state moves
    std::vector<int> all_commits;   // the same as built (commit clears `commits`)
    // Which generator each item and jump item (in their device order) reads
    // and writes, so retarget() can point the work list at other state
    // slots: item_io bit 0 = state_in is the generator's window (else a
    // jumped start state), bit 1 = state_out is the generator's next window.
    std::vector<int32_t> item_gen, jitem_gen;
    std::vector<uint8_t> item_io;
    size_t o_items = 0, o_jitems = 0;  // their offsets in the workspace
    // Matrix-core jumps: groups ordered by phase, items by (row tiles,
    // phase); per phase (0, 1, 2 = merged) the group range and the item
    // range of each row-tile count.  (A group holds all of a generator's
    // jumps and counts as phase 0.)
    std::vector<MJumpGroup> mgroups;
    std::vector<MJumpItem> mitems;
    std::vector<int32_t> mgroup_gen;
    struct MRange {
        int g0 = 0, g1 = 0;
        int i0[kMjRowTiles + 1] = {}, i1[kMjRowTiles + 1] = {};
    } mrange[3];
    size_t o_mgroups = 0;
    MJumpGroup *d_mgroups = nullptr;
    MJumpItem *d_mitems = nullptr;
    int mjumps = 0;  // jumps run on the matrix cores
    hipEvent_t ev_upload = nullptr;    // the last upload from h_stage
    // Device copies: one grow-only workspace [items | segs | polys | masks |
    // jump items], the jumped-state arena and the work-queue counters.
    char *d_work = nullptr;
    size_t work_cap = 0;
    char *h_stage = nullptr;  // pinned staging copy of the workspace
    size_t stage_cap = 0;
    char *d_jstates = nullptr;
    size_t jstates_cap = 0;
    uint32_t *d_queue = nullptr;
    Item *d_items = nullptr;
    JumpItem *d_jitems = nullptr;
    uint64_t segment_used = 0;
    uint64_t numbers = 0;  // numbers generated or stepped over
    int jumps = 0;         // cells that start with a jump
    double plan_ms = 0;
    bool has_phase1 = false;  // the small cells form launches of their own
    // Multi-round grid (or caller-sized cells): default workgroups per CU
    // of multi-wave layouts divided by their waves.
    bool waves_scaled = false;
    bool one_round_grid = false;  // cells from the step-time model, one round
    bool overlap = false;
    // Batched jumps are split into partial jumps (chunk ranges, XORed when
    // the generation kernel loads the start window) until a jump launch has
    // about this many waves per CU: a jump launch is one round of waves, so
    // its duration is one wave's share of the chunks.
    int jump_waves_per_cu = 16;
    // Most partial jumps one batched jump is split into (each reads and
    // XORs its partial window when the generation kernel loads the cell).
    // 16 measured best: one MTGP32-11213 stream (1017 jumps) 0.195 ms per
    // jump launch against 0.25 ms at 8; batches unchanged.
    int jump_max_parts = 16;     // launch() uses the two-phase order
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

    void build(Context &c, const std::vector<GenRef> &gens, const std::vector<mtgp_request_t> &reqs,
               const PlanOptions &opt);
    void build(Context &c, const std::vector<GenRef> &gens, const std::vector<mtgp_request_t> &reqs,
               uint64_t segment) {
        PlanOptions o;
        o.segment = segment;
        build(c, gens, reqs, o);
    }
    // Points the work list at new generator windows (gens[g].state_in /
    // state_out): the same cells, jump polynomials and outputs, continuing
    // from other states.  Patches the items' and jump items' state addresses
    // and uploads them on `s`; no polynomial work.
    void retarget(const std::vector<GenRef> &gens, hipStream_t s);
    // Waits until the last upload from the pinned staging copy has run, so
    // the copy can be rewritten.
    void wait_upload();
    // Sizes the grow-only buffers up front (a plan that is rebuilt often
    // then never frees, which would synchronise the device).
    void reserve(size_t work_bytes, size_t jstates_bytes);
    void launch(hipStream_t s);
    void launch_part(int part, hipStream_t s);  // 0 jumps, 1 generation
    void launch_phase(int part, int phase, hipStream_t s);
    void run_launches(int part, int phase, hipStream_t s);  // phase -1: merged
    bool two_phase() const { return overlap && has_phase1; }
    void ensure_side();  // the side stream and fork/join events
    void set_overlap(bool on) {
        drop_graph();
        overlap = on;
    }
    // Overrides the step layout and/or the resident workgroups per CU of
    // every launch (-1 keeps the current value).
    void configure(int layout, int blocks_per_cu);
    // Records the plan's launches in a hipGraph (on s, not the null stream)
    // and instantiates it; launch_graph replays it.  configure drops it.
    void capture(hipStream_t s);
    void launch_graph(hipStream_t s);
    void drop_graph();
    // Times the generation launches for every step layout and resident
    // workgroups per CU (after one jump launch), configures the fastest and
    // makes it the default of its classes.  Returns the best time in ms.
    float autotune(hipStream_t s, int reps, int *layout, int *bpc);
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    void free_device();
    Plan() = default;
    Plan(const Plan &) = delete;
    Plan &operator=(const Plan &) = delete;
    ~Plan();
};

}  // namespace mtgp

// A device generator set (mtgp_genset_t): two state slots per generator;
// `cur` names the slot holding the current window.  Every change of a
// generator's state (init, set_state, a jump, substreams, a plan's commit,
// a served batch) bumps its `epoch`; a plan records the epochs it was built
// (or rebased) against and refuses to commit over a newer state.
struct mtgp_genset {
    mtgp::Context *ctx;
    std::vector<int> bits, mexp, slot, n, cur;
    std::vector<uint64_t> epoch;
    void set_cur(int g, int which) {
        cur[(size_t)g] = which;
        epoch[(size_t)g]++;
    }
    // Parameter record of each generator (the one of its word size is used).
    std::vector<mtgp32_params_fast_t> p32;
    std::vector<mtgp64_params_fast_t> p64;
    std::vector<size_t> off;  // byte offset of slot pair in d_states
    char *d_states = nullptr;
    // Device work of mtgp_genset_substreams, kept for the next call with the
    // same jump polynomial: [base window | masks | jump items | queue
    // counters | partial windows].
    struct Substreams {
        std::string key;  // jump string the work list was built for
        char *d = nullptr;
        size_t cap = 0;
        std::vector<int> first;  // jump launches: items [first[k], first[k+1])
        int combine_first = 0, combine_count = 0;
        size_t o_items = 0, o_queue = 0;
        // Matrix-core variant: one MJumpGroup of every substream from the
        // base, its items by row-tile count [mi_first[rt], mi_first[rt] +
        // mi_count[rt]) at o_mitems.
        bool matrix = false;
        size_t o_mgroup = 0, o_mitems = 0;
        int mi_first[mtgp::kMjRowTiles + 1] = {}, mi_count[mtgp::kMjRowTiles + 1] = {};
    } sub;
    // mtgp_genset_generate's plan, reused so that repeated calls allocate
    // nothing once their sizes have been seen (calls on one set serialise).
    // A call whose requests equal the last call's, with no planner default
    // changed since (gen_epoch), re-points the built plan at the
    // generators' current windows instead of planning again: its cells,
    // jumps and output addresses depend only on the requests.
    std::mutex gen_mu;
    std::unique_ptr<mtgp::Plan> gen_plan;
    std::vector<mtgp_request_t> gen_reqs;
    uint64_t gen_epoch = ~0ull;
    ~mtgp_genset() {
        gen_plan.reset();
        if (d_states)
            (void)hipFree(d_states);
        if (sub.d)
            (void)hipFree(sub.d);
    }
    void *state(int g, int which) {
        return d_states + off[(size_t)g] + (size_t)which * (size_t)n[(size_t)g] * (size_t)(bits[(size_t)g] / 8);
    }
    std::vector<mtgp::GenRef> refs() {
        std::vector<mtgp::GenRef> r(bits.size());
        for (size_t g = 0; g < bits.size(); g++) {
            r[g].bits = bits[g];
            r[g].mexp = mexp[g];
            r[g].n = n[g];
            r[g].slot = slot[g];
            r[g].state_in = state((int)g, cur[g]);
            r[g].state_out = state((int)g, cur[g] ^ 1);
        }
        return r;
    }
};


