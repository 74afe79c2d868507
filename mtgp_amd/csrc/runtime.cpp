// runtime.cpp -- device context, jump-polynomial service, batched planner
// and the C entry points of mtgp_hip.h / mtgp*-fast-jump.h.
#include <hip/hip_runtime_api.h>

#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "mtgp32-fast.h"
#include "mtgp32-fast-jump.h"
#include "mtgp64-fast.h"
#include "mtgp64-fast-jump.h"
#include "mtgp_hip.h"

#include "common.hpp"
#include "gf2x.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace mtgp {

Error::Error(const std::string &m) : std::runtime_error(m) {}

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------- context

namespace {
std::mutex g_ctx_mu;
std::map<int, std::unique_ptr<Context>> g_ctx;
int g_device = -1;
}  // namespace

Context &context() {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    int dev = g_device;
    if (dev < 0) {
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (count <= 0)
            throw Error("no GPU visible to HIP");
        hip_check(hipGetDevice(&dev), "hipGetDevice");
    }
    auto it = g_ctx.find(dev);
    if (it != g_ctx.end()) {
        hip_check(hipSetDevice(dev), "hipSetDevice");
        return *it->second;
    }
    hip_check(hipSetDevice(dev), "hipSetDevice");
    std::unique_ptr<Context> c(new Context());
    c->device = dev;
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    c->cus = prop.multiProcessorCount;
    c->hbm_bytes = prop.totalGlobalMem;
    c->arch = prop.gcnArchName;
    // Parameter table: the 768 reference sets, then room for custom sets.
    c->param_capacity = kNumSets + 1024;
    c->customs.reserve((size_t)(c->param_capacity - kNumSets));
    hip_check(hipMalloc((void **)&c->d_params, sizeof(DevParam) * (size_t)c->param_capacity), "hipMalloc params");
    std::vector<DevParam> host(kNumSets);
    for (int k = 0; k < kNumSets; k++)
        host[(size_t)k] = host_param(k);
    hip_check(hipMemcpy(c->d_params, host.data(), sizeof(DevParam) * host.size(), hipMemcpyHostToDevice),
              "hipMemcpy params");
    c->n_params = kNumSets;
    for (int b = 0; b < 2; b++)
        for (int cls = 0; cls < kClasses; cls++)
            for (int lay = 0; lay < kLayouts; lay++) {
                if (b == 0 && cls == 3)
                    continue;
                // (a layout the class lacks runs as layout 1)
                const int kl = layout_supported(b ? 64 : 32, cls, lay) ? lay : 1;
                int nb = 0;
                hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel_fn(b ? 64 : 32, cls, kl),
                                                                       kernel_threads(b ? 64 : 32, cls, kl), 0),
                          "hipOccupancyMaxActiveBlocksPerMultiprocessor");
                c->blocks_per_cu[b][cls][lay] = nb > 0 ? nb : 1;
            }
    for (int b = 0; b < 2; b++)
        for (int cls = 0; cls < kClasses; cls++) {
            if (b == 0 && cls == 3)
                continue;
            int nb = 0;
            hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, jump_kernel_fn(b ? 64 : 32, cls), 64, 0),
                      "hipOccupancyMaxActiveBlocksPerMultiprocessor");
            c->jump_blocks_per_cu[b][cls] = nb > 0 ? nb : 1;
        }
    if (const char *e = getenv("MTGP_BLOCKS_PER_CU")) {
        const int v = atoi(e);
        if (v > 0) {
            c->default_bpc = v;
            for (int b = 0; b < 2; b++)
                for (int cls = 0; cls < kClasses; cls++)
                    c->bpc[b][cls] = v;
        }
    }
    if (const char *e = getenv("MTGP_SEGMENT_SPLIT")) {
        const double v = atof(e);
        if (v >= 0 && v < 1)
            c->segment_split = v;
    }
    if (const char *e = getenv("MTGP_OVERLAP"))
        c->overlap = atoi(e) != 0;
    if (const char *e = getenv("MTGP_SMALL_ROUNDS")) {
        const double v = atof(e);
        if (v >= 0.1 && v <= 8)
            c->small_rounds = v;
    }
    if (const char *e = getenv("MTGP_TWO_SIZE_RHO"))
        c->two_size_rho = atof(e);
    if (const char *e = getenv("MTGP_JUMP_FLOOR_MS"))
        c->jump_floor_ms = std::max(0.0, atof(e));
    if (const char *e = getenv("MTGP_JUMP_MAX_PARTS")) {
        const int v = atoi(e);
        if (v > 0 && v <= 64)
            c->jump_max_parts = v;
    }
    if (const char *e = getenv("MTGP_JUMP_MFMA_MIN")) {
        const int v = atoi(e);
        if (v >= 0)
            c->jump_mfma_min = v;
    }
    if (const char *e = getenv("MTGP_JUMP_WAVES")) {
        const int v = atoi(e);
        if (v > 0 && v <= 64)
            c->jump_waves_per_cu = v;
    }
    if (const char *e = getenv("MTGP_SEGMENT_ROUNDS")) {
        const double v = atof(e);
        if (v >= 0.25 && v <= 64)
            c->segment_rounds = v;
    }
    if (const char *e = getenv("MTGP_LAYOUT")) {
        const int lay = atoi(e);
        if (lay >= 0 && lay < kLayouts)
            for (int b = 0; b < 2; b++)
                for (int cls = 0; cls < 3; cls++)  // (class 3 is lane-strided only)
                    if (layout_supported(b ? 64 : 32, cls, lay))
                        c->layout[b][cls] = c->layout_one_round[b][cls] = lay;
    }
    Context &ref = *c;
    g_ctx[dev] = std::move(c);
    return ref;
}

void select_device(int dev) {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    g_device = dev;
}

size_t Context::lds_reserve(const void *fn, int threads, int per_cu) {
    // Off unless asked for: measured neutral at 16 per CU.
    const char *e = getenv("MTGP_LDS_CAP");
    if (!e || atoi(e) == 0)
        return 0;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_pair(fn, per_cu);
    auto it = reserve_cache.find(key);
    if (it != reserve_cache.end())
        return it->second;
    int base = 0;
    hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&base, fn, threads, 0),
              "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    size_t dyn = 0;
    if (base > per_cu) {
        hipFuncAttributes attr;
        hip_check(hipFuncGetAttributes(&attr, fn), "hipFuncGetAttributes");
        int lds_cu = 0;
        hip_check(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device),
                  "hipDeviceGetAttribute");
        // Largest reservation that still allows per_cu workgroups.
        size_t want = (size_t)lds_cu / (size_t)per_cu;
        dyn = want > attr.sharedSizeBytes ? want - attr.sharedSizeBytes : 0;
        dyn &= ~(size_t)255;
        for (int tries = 0; tries < 64; tries++) {
            int nb = 0;
            hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 64, dyn),
                      "hipOccupancyMaxActiveBlocksPerMultiprocessor");
            if (nb <= per_cu || dyn == 0)
                break;
            dyn += 256;
        }
    }
    reserve_cache[key] = dyn;
    return dyn;
}

// Parameter slot of an arbitrary (host) parameter record: a table set when
// it equals one, else a custom slot uploaded on first use.
int Context::slot_of(const DevParam &rec, const unsigned char *sha1) {
    // A device step computes 256 words at once, or 128 for the narrow class
    // (MTGP64 states of N <= 176, like the dynamic-creator sets), which needs
    // N - pos >= 256 (128).
    if (kernel_class(rec.bits, rec.n, rec.pos) < 0)
        throw Error("parameter set with N = " + std::to_string(rec.n) + ", N - pos = " +
                    std::to_string(rec.n - rec.pos) + " fits no device kernel (MTGP" + std::to_string(rec.bits) +
                    "-" + std::to_string(rec.mexp) + ")");
    for (int k = 0; k < kNumSets; k++)
        if (memcmp(&host_param(k), &rec, sizeof(DevParam)) == 0)
            return k;
    std::lock_guard<std::mutex> lk(mu);
    for (size_t k = 0; k < customs.size(); k++)
        if (memcmp(&customs[k].rec, &rec, sizeof(DevParam)) == 0)
            return kNumSets + (int)k;
    if (n_params >= param_capacity)
        throw Error("too many custom parameter sets");
    hip_check(hipMemcpy(d_params + n_params, &rec, sizeof(DevParam), hipMemcpyHostToDevice), "hipMemcpy param");
    Custom cst;
    cst.rec = rec;
    memcpy(cst.sha1, sha1, 20);
    customs.push_back(cst);
    return n_params++;
}

int Context::class_of(int slot) const {
    const DevParam &d = param(slot);
    return kernel_class(d.bits, d.n, d.pos);
}

int Context::layouts_ok(int slot) const {
    const DevParam &d = param(slot);
    const int cls = kernel_class(d.bits, d.n, d.pos);
    int mask = 0;
    for (int lay = 0; lay < kLayouts; lay++)
        if (lay < 2 || (layout_supported(d.bits, cls, lay) && d.n - d.pos >= layout_step_words(d.bits, cls, lay)))
            mask |= 1 << lay;
    return mask;
}

int Context::layout_within(int lay, int mask) {
    while (lay >= 2 && !(mask >> lay & 1))
        lay--;
    return lay;
}

int Context::layout_for(int bits, int cls, int mask, bool one_round) const {
    return layout_within(one_round ? layout_one_round[bits == 64][cls] : layout[bits == 64][cls], mask);
}

const DevParam &Context::param(int slot) const {
    return slot < kNumSets ? host_param(slot) : customs[(size_t)(slot - kNumSets)].rec;
}

void Context::sha1_of(int slot, unsigned char out[20]) const {
    if (slot < kNumSets) {
        const int t = slot / kSetsPerTable, i = slot % kSetsPerTable;
        static mtgp32_params_fast_t *const t32[3] = {mtgp32_params_fast_11213, mtgp32_params_fast_23209,
                                                     mtgp32_params_fast_44497};
        static mtgp64_params_fast_t *const t64[3] = {mtgp64_params_fast_23209, mtgp64_params_fast_44497,
                                                     mtgp64_params_fast_110503};
        memcpy(out, t < 3 ? t32[t][i].poly_sha1 : t64[t - 3][i].poly_sha1, 20);
    } else {
        memcpy(out, customs[(size_t)(slot - kNumSets)].sha1, 20);
    }
}

DevParam dev_param_of(const mtgp32_params_fast_t &p) {
    DevParam d{};
    d.bits = 32;
    d.mexp = p.mexp;
    d.n = p.mexp / 32 + 1;
    d.pos = p.pos;
    d.sh1 = p.sh1;
    d.sh2 = p.sh2;
    d.mask_lo = p.mask;
    for (int j = 0; j < 16; j++) {
        d.tbl[j] = p.tbl[j];
        d.tmp[j] = p.tmp_tbl[j];
        d.flt[j] = p.flt_tmp_tbl[j];
    }
    return d;
}

DevParam dev_param_of(const mtgp64_params_fast_t &p) {
    DevParam d{};
    d.bits = 64;
    d.mexp = p.mexp;
    d.n = p.mexp / 64 + 1;
    d.pos = p.pos;
    d.sh1 = p.sh1;
    d.sh2 = p.sh2;
    d.mask_lo = (uint32_t)p.mask;
    d.mask_hi = (uint32_t)(p.mask >> 32);
    for (int j = 0; j < 16; j++) {
        d.tbl[j] = (uint32_t)(p.tbl[j] >> 32);
        d.tmp[j] = (uint32_t)(p.tmp_tbl[j] >> 32);
        d.flt[j] = (uint32_t)(p.dbl_tmp_tbl[j] >> 32);
    }
    return d;
}

// The kernels keep only the high halves of MTGP64 tables; a custom set
// with non-zero low halves cannot run on them.
static void check64(const mtgp64_params_fast_t &p) {
    for (int j = 0; j < 16; j++)
        if ((uint32_t)p.tbl[j] || (uint32_t)p.tmp_tbl[j] || (uint32_t)p.dbl_tmp_tbl[j])
            throw Error("mtgp64 parameter tables with non-zero low halves are not supported");
}

// ----------------------------------------------- characteristic polynomials

namespace {

// Host stream of `count` output LSBs from the window of `rec` (seed 1),
// the sequence calc_characteristic feeds to MinPolySeq.
std::vector<uint8_t> lsb_sequence(const DevParam &d, long count) {
    std::vector<uint8_t> bits((size_t)count);
    const int n = d.n, pos = d.pos;
    if (d.bits == 32) {
        std::vector<uint32_t> w((size_t)(n + count));
        mtgp32_params_fast_t p{};
        p.mexp = d.mexp;
        for (int j = 0; j < 16; j++)
            p.tbl[j] = d.tbl[j];
        mtgp32_init_state(w.data(), &p, 1);
        for (long k = 0; k < count; k++) {
            uint32_t x = (w[(size_t)k] & d.mask_lo) ^ w[(size_t)k + 1];
            x ^= x << d.sh1;
            uint32_t y = x ^ (w[(size_t)(k + pos)] >> d.sh2);
            y ^= d.tbl[y & 15];
            w[(size_t)(n + k)] = y;
            uint32_t t = w[(size_t)(k + pos - 1)];
            t ^= t >> 16;
            t ^= t >> 8;
            bits[(size_t)k] = (uint8_t)((y ^ d.tmp[t & 15]) & 1);
        }
    } else {
        std::vector<uint64_t> w((size_t)(n + count));
        mtgp64_params_fast_t p{};
        p.mexp = d.mexp;
        for (int j = 0; j < 16; j++)
            p.tbl[j] = (uint64_t)d.tbl[j] << 32;
        mtgp64_init_state(w.data(), &p, 1);
        const uint64_t mask = ((uint64_t)d.mask_hi << 32) | d.mask_lo;
        for (long k = 0; k < count; k++) {
            uint64_t x = (w[(size_t)k] & mask) ^ w[(size_t)k + 1];
            uint32_t xh = (uint32_t)(x >> 32), xl = (uint32_t)x;
            const uint64_t y = w[(size_t)(k + pos)];
            xh ^= xh << d.sh1;
            xl ^= xl << d.sh1;
            const uint32_t nh = xl ^ ((uint32_t)(y >> 32) >> d.sh2);
            const uint32_t nl = xh ^ ((uint32_t)y >> d.sh2);
            w[(size_t)(n + k)] = ((uint64_t)(nh ^ d.tbl[nl & 15]) << 32) | nl;
            // the tempering table's low halves are zero: the LSB is nl's
            bits[(size_t)k] = (uint8_t)(nl & 1);
        }
    }
    return bits;
}

std::mutex g_poly_mu;
std::unordered_map<std::string, std::shared_ptr<gf2::Modulus>> g_mod;  // bits:mexp:sha1 -> P

}  // namespace

std::shared_ptr<gf2::Modulus> modulus_for(const DevParam &d, const unsigned char *sha1) {
    char key[64];
    snprintf(key, sizeof key, "%d:%d:", d.bits, d.mexp);
    std::string k(key);
    for (int i = 0; i < 20; i++) {
        snprintf(key, sizeof key, "%02x", sha1[i]);
        k += key;
    }
    {
        std::lock_guard<std::mutex> lk(g_poly_mu);
        auto it = g_mod.find(k);
        if (it != g_mod.end())
            return it->second;
    }
    gf2::Poly P = gf2::berlekamp_massey(lsb_sequence(d, 2L * d.mexp));
    if (gf2::degree(P) != d.mexp)
        throw Error("characteristic polynomial has degree " + std::to_string(gf2::degree(P)) + ", expected " +
                    std::to_string(d.mexp));
    const std::string s = gf2::coeff_string(P);
    unsigned char h[20];
    gf2::sha1(s.data(), s.size(), h);
    if (memcmp(h, sha1, 20) != 0)
        throw Error("characteristic polynomial does not match the parameter set's poly_sha1");
    auto m = std::make_shared<gf2::Modulus>(P);
    std::lock_guard<std::mutex> lk(g_poly_mu);
    g_mod[k] = m;
    return m;
}

std::shared_ptr<gf2::Modulus> modulus_for(Context &c, int slot) {
    unsigned char sha[20];
    c.sha1_of(slot, sha);
    return modulus_for(c.param(slot), sha);
}

// ---------------------------------------------------------------- planning

namespace {

int elem_bytes(int bits) { return bits / 8; }

// Persistent pool of up to 15 worker threads (plus the caller) for
// parallel_for; a call hands out indices from a shared counter.  Calls from
// several threads at once are serialised; a call from inside a task runs
// inline.
thread_local bool t_in_pool = false;

class Pool {
  public:
    static Pool &get() {
        static Pool p;
        return p;
    }
    void run(int n, const std::function<void(int)> &f) {
        std::lock_guard<std::mutex> call(call_mu_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &f;
            n_ = n;
            next_ = 0;
            active_ = (int)workers_.size();
            err_.clear();
            gen_++;
        }
        cv_.notify_all();
        t_in_pool = true;
        work();
        t_in_pool = false;
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return active_ == 0; });
        fn_ = nullptr;
        if (!err_.empty())
            throw Error(err_);
    }

    // Threads a call runs on (the workers and the caller).
    int threads() const { return (int)workers_.size() + 1; }

  private:
    Pool() {
        const int hw = (int)std::thread::hardware_concurrency();
        const int nt = std::max(0, std::min(hw > 0 ? hw : 1, 16) - 1);
        for (int t = 0; t < nt; t++)
            workers_.emplace_back([this] { loop(); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &w : workers_)
            w.join();
    }
    void work() {
        for (;;) {
            const int i = next_++;
            if (i >= n_)
                return;
            try {
                (*fn_)(i);
            } catch (const std::exception &e) {
                std::lock_guard<std::mutex> lk(mu_);
                err_ = e.what();
                next_ = n_;
            }
        }
    }
    void loop() {
        t_in_pool = true;
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_)
                    return;
                seen = gen_;
            }
            work();
            std::lock_guard<std::mutex> lk(mu_);
            if (--active_ == 0)
                done_cv_.notify_all();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex call_mu_, mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int)> *fn_ = nullptr;
    std::atomic<int> next_{0};
    int n_ = 0, active_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
    std::string err_;
};

// Parallel map over [0, n) (inline for n <= 1).
template <class F>
void parallel_for(int n, F f) {
    if (n <= 1 || t_in_pool) {
        for (int i = 0; i < n; i++)
            f(i);
        return;
    }
    const std::function<void(int)> fn = f;
    Pool::get().run(n, fn);
}

}  // namespace

std::vector<uint16_t> jump_mask16(const std::vector<uint32_t> &w) {
    const size_t ncoeff = 32 * w.size();
    const int nchunks = std::max(1, (int)((ncoeff + kJumpChunk - 1) / kJumpChunk));
    std::vector<uint16_t> m((size_t)nchunks * 64, 0);
    for (int c = 0; c < nchunks; c++) {
        // Row q of the chunk = coefficients c*1024 + 64q .. +63; the mask of
        // residue rho is column rho of this 16 x 64 bit matrix.
        uint64_t rows[16];
        for (int q = 0; q < 16; q++) {
            const size_t i = (size_t)c * 32 + 2 * (size_t)q;
            const uint64_t lo = i < w.size() ? w[i] : 0u, hi = i + 1 < w.size() ? w[i + 1] : 0u;
            rows[q] = lo | hi << 32;
        }
        uint16_t *mc = m.data() + (size_t)c * 64;
        for (int rho = 0; rho < 64; rho++) {
            uint32_t v = 0;
            for (int q = 0; q < 16; q++)
                v |= (uint32_t)((rows[q] >> rho) & 1u) << q;
            mc[rho] = (uint16_t)v;
        }
    }
    return m;
}

std::vector<uint32_t> words_from_mask16(const std::vector<uint16_t> &m, int nwords) {
    std::vector<uint32_t> w((size_t)nwords, 0);
    const size_t nchunks = m.size() / 64;
    for (size_t c = 0; c < nchunks; c++) {
        uint64_t rows[16] = {};
        const uint16_t *mc = m.data() + c * 64;
        for (int rho = 0; rho < 64; rho++)
            for (int q = 0; q < 16; q++)
                rows[q] |= (uint64_t)((mc[rho] >> q) & 1u) << rho;
        for (int q = 0; q < 16; q++) {
            const size_t i = c * 32 + 2 * (size_t)q;
            if (i < w.size())
                w[i] = (uint32_t)rows[q];
            if (i + 1 < w.size())
                w[i + 1] = (uint32_t)(rows[q] >> 32);
        }
    }
    return w;
}

int append_jump_masks16(const std::vector<const std::vector<uint16_t> *> &ms, std::vector<uint64_t> &out) {
    size_t n = 64;
    for (const auto *m : ms)
        n = std::max(n, m->size());
    const size_t base = out.size();
    out.resize(base + n, 0);
    uint64_t *o = out.data() + base;
    for (size_t b = 0; b < ms.size(); b++) {
        const std::vector<uint16_t> &m = *ms[b];
        for (size_t i = 0; i < m.size(); i++)
            o[i] |= (uint64_t)m[i] << (16 * b);
    }
    return (int)(n / 64);
}

int append_jump_masks(const std::vector<const std::vector<uint32_t> *> &polys, std::vector<uint64_t> &out) {
    std::vector<std::vector<uint16_t>> ms;
    for (const auto *w : polys)
        ms.push_back(jump_mask16(*w));
    std::vector<const std::vector<uint16_t> *> ps;
    for (const auto &m : ms)
        ps.push_back(&m);
    return append_jump_masks16(ps, out);
}

uint64_t auto_segment(const Context &c, uint64_t total_numbers, int n_requests, int slots) {
    // Segments of one launch run side by side, so a launch takes about
    // (its work / resident slots) plus the tail of the last round.  About
    // `segment_rounds` equal segments per slot, less room for the one partial
    // segment each request ends with.
    const double want = c.segment_rounds * (double)slots - (double)n_requests;
    const uint64_t usable = (uint64_t)std::max(1.0, want);
    uint64_t s = (total_numbers + usable - 1) / usable;
    s = std::max<uint64_t>(s, 4096);
    return (s + 255) & ~255ull;
}

namespace {
gf2::Poly poly_of_words32(const std::vector<uint32_t> &w) {
    gf2::Poly p((w.size() + 1) / 2, 0);
    for (size_t i = 0; i < w.size(); i++)
        p[i / 2] |= (uint64_t)w[i] << (32 * (i % 2));
    return p;
}
}  // namespace

JumpCache::JumpCache(Context &c, const std::vector<int> &slots, uint64_t cell, int kmax, bool far) : cell_(cell) {
    if (cell == 0 || kmax < 1)
        throw Error("jump cache needs a cell size and kmax >= 1");
    B_ = (uint64_t)kmax + 1;
    constexpr int kT2 = 64;
    std::vector<int> keys(slots);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    std::vector<Entry *> list;
    for (int s : keys)
        list.push_back(&entries_[s]);
    // One task per (slot, block of table entries): the table is a chain of
    // products, so each block starts from its own x^(k0 cell).
    constexpr uint64_t kBlock = 64;
    const uint64_t nblocks = (B_ + kBlock - 1) / kBlock;
    parallel_for((int)list.size(), [&](int i) {
        Entry &e = *list[(size_t)i];
        e.mod = modulus_for(c, keys[(size_t)i]);
        e.words32 = (int)((e.mod->deg() + 31) / 32);
        e.pow2.push_back(e.mod->xpow(cell));
        e.t1.resize(B_);
        e.t1w.resize(B_);
        if (!far)
            return;
        for (int k = 1; k < 48; k++)
            e.pow2.push_back(e.mod->sqrmod(e.pow2.back()));
        const gf2::Poly big = e.mod->xpow(std::vector<uint64_t>{B_ * cell});
        e.t2p.resize(kT2);
        e.t2p[0] = gf2::Poly{1};
        e.t2p[1] = big;
        for (int j = 2; j < kT2; j++)
            e.t2p[(size_t)j] = e.mod->mulmod(e.t2p[(size_t)j - 1], big);
    });
    parallel_for((int)(list.size() * nblocks), [&](int t) {
        Entry &e = *list[(size_t)t / nblocks];
        const uint64_t k0 = (uint64_t)t % nblocks * kBlock, k1 = std::min(B_, k0 + kBlock);
        gf2::Poly cur = k0 ? e.mod->xpow(k0 * cell) : gf2::Poly{1};
        for (uint64_t k = k0; k < k1; k++) {
            if (k > 0) {
                cur = k == k0 ? cur : e.mod->mulmod(cur, e.pow2[0]);
                e.t1w[k] = gf2::to_words32(cur, e.words32);
                e.t1[k] = jump_mask16(e.t1w[k]);
            }
        }
    });
}

namespace {
std::mutex g_cache_mu;
std::map<std::tuple<int, uint64_t, int, std::vector<int>>, std::shared_ptr<JumpCache>> g_caches;
}  // namespace

std::shared_ptr<JumpCache> JumpCache::shared(Context &c, const std::vector<int> &slots, uint64_t cell, int kmax,
                                             bool far) {
    std::vector<int> keys(slots);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    auto key = std::make_tuple(c.device, cell, far ? kmax : -kmax, keys);
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto it = g_caches.find(key);
    if (it != g_caches.end())
        return it->second;
    // at most kMaxCaches kept: drop the ones nobody else holds
    constexpr size_t kMaxCaches = 32;
    if (g_caches.size() >= kMaxCaches)
        for (auto e = g_caches.begin(); e != g_caches.end();)
            e = e->second.use_count() == 1 ? g_caches.erase(e) : std::next(e);
    auto p = std::make_shared<JumpCache>(c, keys, cell, kmax, far);
    g_caches[key] = p;
    return p;
}

bool JumpCache::get_masks(int slot, uint64_t pos, std::vector<uint16_t> &m16, bool cheap) const {
    if (pos == 0 || pos % cell_)
        return false;
    auto it = entries_.find(slot);
    if (it == entries_.end())
        return false;
    const Entry &e = it->second;
    const uint64_t k = pos / cell_;
    if (k < B_) {
        m16 = e.t1[(size_t)k];
        return true;
    }
    if (cheap)
        return false;
    const uint64_t j = k / B_, r = k % B_;
    gf2::Poly acc;
    if (j < e.t2p.size()) {
        acc = r ? e.mod->mulmod(e.t2p[(size_t)j], poly_of_words32(words_from_mask16(e.t1[(size_t)r], e.words32)))
                : e.t2p[(size_t)j];
    } else {
        if (k >> e.pow2.size())
            return false;
        for (size_t i = 0; i < e.pow2.size(); i++)
            if (k >> i & 1)
                acc = acc.empty() ? e.pow2[i] : e.mod->mulmod(acc, e.pow2[i]);
    }
    m16 = jump_mask16(gf2::to_words32(acc, e.words32));
    return true;
}

const std::vector<uint32_t> *JumpCache::table_words(int slot, uint64_t pos) const {
    if (pos == 0 || pos % cell_)
        return nullptr;
    auto it = entries_.find(slot);
    if (it == entries_.end() || pos / cell_ >= B_)
        return nullptr;
    return &it->second.t1w[(size_t)(pos / cell_)];
}

bool JumpCache::get(int slot, uint64_t pos, std::vector<uint32_t> &words, bool cheap) const {
    std::vector<uint16_t> m16;
    if (!get_masks(slot, pos, m16, cheap))
        return false;
    auto it = entries_.find(slot);
    words = words_from_mask16(m16, it->second.words32);
    return true;
}

namespace {
// Grow-only device / pinned buffers (a plan's buffers are reused by later
// builds; freeing synchronises the device, so they grow rarely: at least
// double, at least 1 MiB).
void grow_device(char *&p, size_t &cap, size_t need, const char *what) {
    if (need <= cap)
        return;
    if (p)
        hip_check(hipFree(p), "hipFree");
    p = nullptr;
    const size_t want = std::max({need, 2 * cap, (size_t)1 << 20});
    cap = 0;
    hip_check(hipMalloc((void **)&p, want), what);
    cap = want;
}
void grow_pinned(char *&p, size_t &cap, size_t need) {
    if (need <= cap)
        return;
    if (p)
        hip_check(hipHostFree(p), "hipHostFree");
    p = nullptr;
    const size_t want = std::max({need, 2 * cap, (size_t)1 << 20});
    cap = 0;
    hip_check(hipHostMalloc((void **)&p, want, hipHostMallocDefault), "hipHostMalloc staging");
    cap = want;
}
size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }
}  // namespace

double jump_cost_numbers(int mexp, int n) { return 0.026 * (double)mexp * (double)n; }

// Plans on a uniform grid with a generator of at least kAutoCacheCells
// cells take their jump polynomials from a process-wide JumpCache of
// kAutoCacheCells * 2^i entries, up to kAutoCacheMaxK.
constexpr int kAutoCacheCells = 64;
constexpr int kAutoCacheMaxK = 2048;
// Cells per block of an incremental polynomial chain on the host pool.
constexpr size_t kChainMin = 32;

GridModel grid_model_of(int bits, int cls, int mexp, int n) {
    GridModel m;
    m.jump_ns = 4.5e-5 * (double)mexp * (double)n * (double)(bits / 32) / 2;
    m.w0 = bits == 64 ? 2.7 : 5.5;
    m.rate = 4700;
    // Step times of 64-set batches (2^24 + 2^24 numbers each) at 4, 8 and
    // 12 cells per CU and of one MTGP32-11213 stream: jump launch time per
    // jump and the generation rate's saturation in waves per CU.
    if (bits == 32 && cls == 0)
        m.jump_ns = 88, m.rate = 4900, m.w0 = 5.0;
    else if (bits == 32 && cls == 1)
        m.jump_ns = 375, m.rate = 3800;
    else if (bits == 32 && cls == 2)
        m.jump_ns = 880, m.rate = 3400;
    else if (bits == 64 && cls == 0)
        m.jump_ns = 380, m.rate = 4750;
    else if (bits == 64 && cls == 1)
        m.jump_ns = 1100, m.rate = 4450;
    return m;
}

bool auto_grid(const Context &c, uint64_t total_numbers, int n_requests, int n_gens, int slots, double jump_cost,
               const GridModel &model, uint64_t &S1, uint64_t &S2, double &frac) {
    const double usable = std::max(1.0, (double)slots - (double)n_requests);
    auto round256 = [](double v) { return std::max<uint64_t>(((uint64_t)v + 255) & ~255ull, 4096); };
    const double rho = jump_cost / std::max(1.0, (double)total_numbers / usable);
    if (rho > c.two_size_rho) {
        // One round of equal cells, w per CU, w minimising the modelled step
        // time (Context::two_size_rho).  w is a multiple of the 4 SIMDs of a
        // CU: an uneven count leaves a CU, or a SIMD, one wave heavier, and
        // its waves finish last (measured: MTGP64-44497 generates no faster
        // at 5 per CU than at 4).
        // With four-wave cells, w = 1..3 (4 to 12 waves per CU) are
        // candidates too (MTGP64-44497: 512 four-wave cells 3.82 TB/s
        // against 3.46 for 1024).
        const int wmax = std::max(1, slots / std::max(1, c.cus));
        const int wstep = model.cell_waves >= 4 ? 1 : 4;
        double best = 0, best_t = 0;
        for (int w = std::min(wstep, wmax); w <= wmax; w += wstep) {
            const double jumps = std::max(0.0, (double)w * c.cus - (double)n_gens);
            const double t_jump = c.jump_floor_ms * 1e6 + jumps * model.jump_ns;
            const double waves = (double)w * model.cell_waves;
            const double t_gen = model.bytes / (model.rate * (1.0 - std::exp(-waves / model.w0)));  // ns
            const double t = t_jump + t_gen;
            if (best == 0 || t < best_t) {
                best = w;
                best_t = t;
            }
        }
        // each generator's last cell is on average half full
        const double cells = std::max(1.0, best * (double)c.cus - 0.5 * (double)n_gens);
        frac = 0;
        S1 = S2 = round256(std::ceil((double)total_numbers / cells));
        return true;
    }
    // Multi-round grids: rounds of slots / cell_waves cells (a large
    // MTGP64-23209 batch in two-wave cells: 4.88 TB/s at 6 per CU against
    // 4.54 at 12 and 4.55 in single-wave cells at 12).
    const int wslots = std::max(1, (int)std::lround((double)slots / std::max(1.0, model.cell_waves_multi)));
    const double wusable = std::max(1.0, (double)wslots - (double)n_requests);
    frac = c.segment_split;
    if (frac <= 0 || frac >= 1) {
        frac = 0;
        S1 = S2 = auto_segment(c, total_numbers, n_requests, wslots);
        return false;
    }
    S1 = round256(std::ceil(frac * (double)total_numbers / wusable));
    S2 = round256(std::ceil((1 - frac) * (double)total_numbers / (wusable * c.small_rounds)));
    return false;
}

void Plan::build(Context &c, const std::vector<GenRef> &gens, const std::vector<mtgp_request_t> &reqs,
                 const PlanOptions &opt) {
    const auto t0 = std::chrono::steady_clock::now();
    ctx = &c;
    // The planner defaults (layouts, workgroups per CU, grid knobs) stay
    // fixed while this plan reads them.
    std::shared_lock<std::shared_mutex> tune_lock(c.tune_mu);
    const int G = (int)gens.size();
    uint64_t all = 0;
    int nonempty = 0;
    // Generators the requests name, in first-use order, with their counts.
    std::vector<int> named;
    std::vector<uint64_t> per((size_t)G, 0);
    std::vector<char> seen((size_t)G, 0);
    for (const auto &r : reqs) {
        if (r.gen < 0 || r.gen >= G)
            throw Error("request for generator " + std::to_string(r.gen) + " outside the set");
        if (r.kind < 0 || r.kind > 4)
            throw Error("unknown output kind " + std::to_string(r.kind));
        if (r.count && !r.out)
            throw Error("request with a NULL output buffer");
        if (r.count && ((uintptr_t)r.out % (uintptr_t)elem_bytes(gens[(size_t)r.gen].bits)))
            throw Error("output buffer not aligned to its element size");
        all += r.count;
        nonempty += r.count > 0;
        per[(size_t)r.gen] += r.count;
        if (!seen[(size_t)r.gen]) {
            seen[(size_t)r.gen] = 1;
            named.push_back(r.gen);
        }
    }
    // Start states: batched jumps unless fused ones are asked for;
    // narrow-class items always need batched start states.
    jump_mode = 1;
    if (opt.jump_mode >= 0)
        jump_mode = opt.jump_mode ? 1 : 0;
    else if (const char *e = getenv("MTGP_JUMP_MODE"))
        jump_mode = atoi(e) ? 1 : 0;
    for (const auto &g : gens)
        if (g.bits == 64 && c.class_of(g.slot) == 3)
            jump_mode = 1;
    // Layouts per launch group: multi-wave cells (layouts 2 and 3) where
    // every generator of the group admits them and start states come from
    // the batched jump launch (the multi-wave kernels have no fused jump).
    int group_ok[7];
    for (int &m : group_ok)
        m = jump_mode == 1 ? (1 << kLayouts) - 1 : 3;
    for (const auto &g : gens)
        group_ok[(g.bits == 64 ? 3 : 0) + c.class_of(g.slot)] &= c.layouts_ok(g.slot);
    // Resident workgroups of the default configuration of the largest class
    // among the plan's generators, and the numbers-weighted mean jump cost.
    int slots = 0;
    double jcost = 0, weight = 0;
    int active = 0;
    GridModel model;
    model.rate = 0;
    model.cell_waves = 0;
    model.cell_waves_multi = 0;
    double w0 = 0;
    for (int g : named) {
        const GenRef &gr = gens[(size_t)g];
        const int b = gr.bits == 64, cls = c.class_of(gr.slot);
        const int lay = c.layout_for(gr.bits, cls, group_ok[3 * b + cls]);
        const int lay1 = c.layout_for(gr.bits, cls, group_ok[3 * b + cls], true);
        slots = std::max(slots, c.cus * std::min(c.bpc[b][cls], c.blocks_per_cu[b][cls][lay]));
        const double nw = (double)per[(size_t)g];
        jcost += nw * jump_cost_numbers(gr.mexp, gr.n);
        const GridModel gm = grid_model_of(gr.bits, cls, gr.mexp, gr.n);
        model.jump_ns += nw * gm.jump_ns;
        model.rate += nw * gm.rate;
        model.bytes += nw * (gr.bits / 8);
        model.cell_waves += nw * layout_waves(lay1);
        model.cell_waves_multi += nw * layout_waves(lay);
        w0 += nw * gm.w0;
        weight += nw;
        active += per[(size_t)g] > 0;
    }
    jcost = weight > 0 ? jcost / weight : 0;
    if (weight > 0) {
        model.jump_ns /= weight;
        model.rate /= weight;
        model.w0 = w0 / weight;
        model.cell_waves /= weight;
        model.cell_waves_multi /= weight;
    } else {
        model.rate = 4700;
        model.cell_waves = model.cell_waves_multi = 1;
    }
    uint64_t S1 = 0, S2 = 0;
    double frac = 0;
    // One round of cells from the step-time model: the generation launch
    // holds them all.  Otherwise (several rounds, or cells the caller
    // sized) multi-wave launches take bpc / waves workgroups per CU.
    bool one_round = false;
    if (opt.cache)
        S1 = S2 = opt.cache->cell();
    else if (opt.segment)
        S1 = S2 = opt.segment;
    else
        one_round = auto_grid(c, all, nonempty, active, slots, jcost, model, S1, S2, frac);
    waves_scaled = !one_round;
    one_round_grid = one_round;
    const uint64_t S = S1;
    // Requests as runs of stream positions per generator.
    struct Run {
        int req;
        uint64_t pos, count;
    };
    std::vector<std::vector<Run>> runs((size_t)G);
    std::vector<uint64_t> total((size_t)G, 0);
    std::vector<char> used((size_t)G, 0);
    for (int ri = 0; ri < (int)reqs.size(); ri++) {
        const auto &r = reqs[(size_t)ri];
        uint64_t &pos = total[(size_t)r.gen];
        used[(size_t)r.gen] = 1;
        pos += r.skip;
        if (r.count)
            runs[(size_t)r.gen].push_back(Run{ri, pos, r.count});
        pos += r.count;
    }
    // Per generator: cells of S1 numbers below B[g] (a multiple of S1), of
    // S2 numbers from B[g] on.
    std::vector<uint64_t> Bg((size_t)G, UINT64_MAX);
    if (frac > 0)
        for (int g = 0; g < G; g++)
            Bg[(size_t)g] = (uint64_t)(frac * (double)total[(size_t)g]) / S1 * S1;
    auto cstart = [&](int g, uint64_t pos) {
        const uint64_t B = Bg[(size_t)g];
        return pos < B ? pos / S1 * S1 : B + (pos - B) / S2 * S2;
    };
    auto cend = [&](int g, uint64_t pos) {
        const uint64_t B = Bg[(size_t)g];
        return pos < B ? (pos / S1 + 1) * S1 : B + ((pos - B) / S2 + 1) * S2;
    };
    // Cells.
    struct Piece {
        int req;  // -1: stepped over
        uint64_t out_off, count;
    };
    struct Cell {
        uint64_t start, end;
        std::vector<Piece> pieces;
    };
    std::vector<std::vector<Cell>> cells((size_t)G);
    for (int g = 0; g < G; g++) {
        if (!used[(size_t)g])
            continue;
        auto &cv = cells[(size_t)g];
        // Continue the current cell at `pos` (stepping over a short gap) or
        // open a new one: at its grid point when the gap to it is short, else
        // exactly at pos.
        auto at = [&](uint64_t pos) {
            if (!cv.empty()) {
                Cell &cur = cv.back();
                if (cstart(g, pos) == cstart(g, cur.start) && pos >= cur.end && pos - cur.end <= opt.discard_max) {
                    if (pos > cur.end)
                        cur.pieces.push_back(Piece{-1, 0, pos - cur.end});
                    cur.end = pos;
                    return;
                }
            }
            const uint64_t g0 = cstart(g, pos);
            uint64_t start = (pos - g0 <= opt.discard_max) ? g0 : pos;
            if (!cv.empty() && start < cv.back().end)
                start = pos;
            Cell cell;
            cell.start = start;
            cell.end = pos;
            if (pos > start)
                cell.pieces.push_back(Piece{-1, 0, pos - start});
            cv.push_back(std::move(cell));
        };
        for (const Run &r : runs[(size_t)g]) {
            uint64_t done = 0;
            while (done < r.count) {
                const uint64_t pos = r.pos + done;
                const uint64_t take = std::min(r.count - done, cend(g, pos) - pos);
                at(pos);
                cv.back().pieces.push_back(Piece{r.req, done, take});
                cv.back().end = pos + take;
                done += take;
            }
        }
        if (cv.empty() || cv.back().end != total[(size_t)g])
            at(total[(size_t)g]);
    }
    // Jump polynomials x^start mod P for cells with start > 0, as
    // coefficient words (fused jumps, mode 0) or in per-residue mask form
    // (batched jumps, mode 1): from the cache when it has them (table
    // entries inline, one-product entries on host threads), else
    // incrementally along each generator on host threads.
    // Launch group of a generator: 0..2 = MTGP32 classes, 3..6 = MTGP64
    // classes 0..3.
    auto group_of = [&](const GenRef &g) { return (g.bits == 64 ? 3 : 0) + c.class_of(g.slot); };
    const bool masks = jump_mode == 1;
    // Phase of a cell: 1 for the small cells of a two-size grid (batched
    // jumps only), else 0.  Phase 1 has launches of its own, which the
    // two-phase launch order runs on a side stream.
    const bool phased = masks && frac > 0;
    auto phase_of = [&](int g, size_t k) {
        return (phased && cells[(size_t)g][k].start >= Bg[(size_t)g]) ? 1 : 0;
    };
    std::vector<std::vector<std::vector<uint32_t>>> polys((size_t)G);
    std::vector<std::vector<std::vector<uint16_t>>> m16((size_t)G);
    // A uniform grid (one cell size, no caller cache) with a generator of
    // many cells -- one long stream, as mtgp32_fill and the torch Generator
    // make -- takes its polynomials x^(k S) from a process-wide cache: a
    // repeated request of the same size then needs no polynomial product,
    // and the first builds the table in blocks on the host pool.
    std::shared_ptr<JumpCache> auto_cache;
    const JumpCache *cache = opt.cache;
    if (!cache && frac == 0 && S > 0) {
        uint64_t kmax = 0;
        std::vector<int> slots;
        for (int g = 0; g < G; g++) {
            const auto &cv = cells[(size_t)g];
            if (cv.size() < (size_t)kAutoCacheCells)
                continue;
            slots.push_back(gens[(size_t)g].slot);
            for (const Cell &cl : cv)
                if (cl.start % S == 0)
                    kmax = std::max<uint64_t>(kmax, cl.start / S);
        }
        if (!slots.empty() && kmax <= (uint64_t)kAutoCacheMaxK) {
            int k = kAutoCacheCells;
            while ((uint64_t)k < kmax)
                k *= 2;
            auto_cache = JumpCache::shared(c, slots, S, k, false);
            cache = auto_cache.get();
        }
    }
    auto cached = [&](int g, size_t k, bool cheap) {
        const int slot = gens[(size_t)g].slot;
        const uint64_t pos = cells[(size_t)g][k].start;
        return masks ? cache->get_masks(slot, pos, m16[(size_t)g][k], cheap)
                     : cache->get(slot, pos, polys[(size_t)g][k], cheap);
    };
    std::vector<int> compute;                      // generators with cells off the cache
    std::vector<std::pair<int, size_t>> composed;  // cells the cache makes with one product
    for (int g = 0; g < G; g++) {
        const auto &cv = cells[(size_t)g];
        polys[(size_t)g].resize(cv.size());
        m16[(size_t)g].resize(cv.size());
        bool need = false;
        for (size_t k = 0; k < cv.size(); k++) {
            if (cv[k].start == 0)
                continue;
            if (cache && cached(g, k, true))
                continue;
            if (cache && cv[k].start % S == 0)
                composed.emplace_back(g, k);
            else
                need = true;
        }
        if (need)
            compute.push_back(g);
    }
    parallel_for((int)composed.size(), [&](int i) {
        (void)cached(composed[(size_t)i].first, composed[(size_t)i].second, false);
    });
    // A cell the cache could not make (its set is not in the cache, or the
    // position is beyond it) is computed below with its generator's others.
    for (const auto &gk : composed)
        if (polys[(size_t)gk.first][gk.second].empty() && m16[(size_t)gk.first][gk.second].empty() &&
            std::find(compute.begin(), compute.end(), gk.first) == compute.end())
            compute.push_back(gk.first);
    // The rest incrementally along each generator (x^b = x^a * x^(b - a)),
    // in blocks of cells so that one long stream is spread over the host
    // pool too: each block starts with its own x^start.
    struct Chain {
        int g;
        std::vector<size_t> ks;
    };
    std::vector<Chain> chains;
    {
        const int want = std::max(1, (2 * Pool::get().threads()) / std::max(1, (int)compute.size()));
        for (int g : compute) {
            const auto &cv = cells[(size_t)g];
            std::vector<size_t> ks;
            for (size_t k = 0; k < cv.size(); k++)
                if (cv[k].start > 0 && polys[(size_t)g][k].empty() && m16[(size_t)g][k].empty())
                    ks.push_back(k);
            // blocks of at least kChainMin cells: a block's first power costs
            // about as much as that many products
            const size_t per = std::max<size_t>(kChainMin, (ks.size() + (size_t)want - 1) / (size_t)want);
            for (size_t i = 0; i < ks.size(); i += per)
                chains.push_back(Chain{g, std::vector<size_t>(ks.begin() + (long)i,
                                                              ks.begin() + (long)std::min(ks.size(), i + per))});
        }
    }
    parallel_for((int)chains.size(), [&](int i) {
        const Chain &ch = chains[(size_t)i];
        const int g = ch.g;
        const auto &cv = cells[(size_t)g];
        auto &out = polys[(size_t)g];
        const GenRef &gr = gens[(size_t)g];
        auto mod = modulus_for(c, gr.slot);
        const int pw = (gr.mexp + 31) / 32;
        std::map<uint64_t, gf2::Poly> step_cache;
        gf2::Poly cur;
        uint64_t cur_off = 0;
        for (size_t k : ch.ks) {
            const uint64_t off = cv[k].start;
            if (cur_off == 0) {
                cur = mod->xpow(off);
            } else {
                const uint64_t delta = off - cur_off;
                auto it = step_cache.find(delta);
                if (it == step_cache.end())
                    it = step_cache.emplace(delta, mod->xpow(delta)).first;
                cur = mod->mulmod(cur, it->second);
            }
            cur_off = off;
            out[k] = gf2::to_words32(cur, pw);
            if (masks)
                m16[(size_t)g][k] = jump_mask16(out[k]);
        }
    });
    // Jump groups (jump_mode 1): up to kJumpBatch cells of one generator.
    struct Group {
        int g, phase;
        std::vector<size_t> cells;
        size_t mask_off;  // byte offset into the mask array
        int nchunks;
    };
    std::vector<Group> groups;
    // Matrix-core groups: all of a generator's jumps (both phases: the
    // two-phase launch runs them with phase 0), one bit-plane pass each.
    struct MGroup {
        int g, phase;
        std::vector<size_t> cells;
    };
    std::vector<MGroup> mg;
    std::vector<uint64_t> mask_host;
    std::vector<uint32_t> poly_host;
    if (masks) {
        for (int g = 0; g < G; g++) {
            if (gens[(size_t)g].bits == 32 && c.jump_mfma_min > 0) {
                std::vector<size_t> need;
                for (size_t k = 0; k < cells[(size_t)g].size(); k++)
                    if (cells[(size_t)g][k].start > 0)
                        need.push_back(k);
                if ((int)need.size() >= c.jump_mfma_min) {
                    mg.push_back(MGroup{g, 0, need});
                    continue;
                }
            }
            for (int ph = 0; ph < 2; ph++) {
                std::vector<size_t> need;
                for (size_t k = 0; k < cells[(size_t)g].size(); k++)
                    if (cells[(size_t)g][k].start > 0 && phase_of(g, k) == ph)
                        need.push_back(k);
                for (size_t i0 = 0; i0 < need.size(); i0 += kJumpBatch) {
                    Group gr;
                    gr.g = g;
                    gr.phase = ph;
                    gr.cells.assign(need.begin() + (long)i0,
                                    need.begin() + (long)std::min(need.size(), i0 + kJumpBatch));
                    std::vector<const std::vector<uint16_t> *> ps;
                    for (size_t k : gr.cells)
                        ps.push_back(&m16[(size_t)g][k]);
                    gr.mask_off = mask_host.size() * 8;
                    gr.nchunks = append_jump_masks16(ps, mask_host);
                    groups.push_back(gr);
                }
            }
        }
    }
    // Split each group's chunks over `parts` waves so the jump launch has
    // about jump_waves_per_cu waves per CU; the partial windows are XORed
    // on load.
    // The split stays within one round of the jump kernel's occupancy (the
    // smallest of the plan's classes): a partial second round would double
    // the launch's duration.
    int parts = 1;
    if (!groups.empty()) {
        const int ng = (int)groups.size();
        int occ = 64;
        for (const Group &gr : groups) {
            const GenRef &ref = gens[(size_t)gr.g];
            occ = std::min(occ, c.jump_blocks_per_cu[ref.bits == 64][c.class_of(ref.slot)]);
        }
        const int want = (c.cus * c.jump_waves_per_cu + ng - 1) / ng;
        const int fit = (c.cus * occ) / ng;
        parts = std::max(1, std::min({want, fit, std::max(1, opt.max_parts > 0 ? opt.max_parts : c.jump_max_parts)}));
    }
    std::vector<int> jgroup, jgen_list;
    jitems.clear();
    size_t arena_bytes = 0;
    // (g, cell) -> (arena offset, parts, part stride)
    std::map<std::pair<int, size_t>, std::tuple<uint64_t, int, uint64_t>> slot_of;
    for (const Group &gr : groups) {
        const GenRef &ref = gens[(size_t)gr.g];
        const size_t sbytes = (size_t)ref.n * (size_t)(ref.bits / 8);
        const size_t nb = gr.cells.size();
        const int P = std::min(parts, gr.nchunks);
        const size_t base = arena_bytes;
        for (size_t b = 0; b < nb; b++)
            slot_of[{gr.g, gr.cells[b]}] = std::make_tuple((uint64_t)(base + b * sbytes), P, (uint64_t)(nb * sbytes));
        for (int q = 0; q < P; q++) {
            JumpItem ji{};
            ji.state_in = (uint64_t)(uintptr_t)ref.state_in;
            ji.masks = gr.mask_off;                         // fixed up below
            ji.states_out = base + (size_t)q * nb * sbytes;  // fixed up below
            ji.nb = (int32_t)nb;
            ji.c0 = (int32_t)((long)gr.nchunks * q / P);
            ji.c1 = (int32_t)((long)gr.nchunks * (q + 1) / P);
            ji.param = ref.slot;
            ji.state_bytes = (int32_t)sbytes;
            jitems.push_back(ji);
            jgroup.push_back(2 * group_of(ref) + gr.phase);
            jgen_list.push_back(gr.g);
        }
        arena_bytes += align256((size_t)P * nb * sbytes);
    }
    // Matrix-core groups: planes and jumped windows in the arena, the
    // coefficient words of every jump transposed ([word][jump]).
    mgroups.clear();
    mitems.clear();
    mgroup_gen.clear();
    for (auto &r : mrange)
        r = MRange{};
    mjumps = 0;
    std::vector<uint32_t> coef_host;
    {
        std::stable_sort(mg.begin(), mg.end(), [](const MGroup &a, const MGroup &b) { return a.phase < b.phase; });
        std::vector<MJumpItem> mit[kMjRowTiles + 1][2];
        for (const MGroup &m : mg) {
            const GenRef &ref = gens[(size_t)m.g];
            const int nj = (int)m.cells.size();
            const int njp = (nj + 31) / 32 * 32;
            const int kpad = mj_kpad(ref.mexp);
            const int wc = kpad / 32, pw = mj_pw(kpad, ref.n);
            const size_t sbytes = (size_t)ref.n * 4;
            MJumpGroup gr{};
            gr.state_in = (uint64_t)(uintptr_t)ref.state_in;
            gr.in_parts = 1;
            gr.planes = arena_bytes;  // fixed up below
            arena_bytes += align256((size_t)pw * 32 * 4);
            gr.out = arena_bytes;  // fixed up below
            gr.out_stride = sbytes;
            for (int b = 0; b < nj; b++)
                slot_of[{m.g, m.cells[(size_t)b]}] = std::make_tuple((uint64_t)(arena_bytes + (size_t)b * sbytes), 1,
                                                                     (uint64_t)0);
            arena_bytes += align256((size_t)nj * sbytes);
            gr.coef = (uint64_t)coef_host.size() * 4;  // fixed up below
            coef_host.resize(coef_host.size() + (size_t)wc * (size_t)njp, 0u);
            uint32_t *cf = coef_host.data() + gr.coef / 4;
            // jump b's words in column b: the cell's own words, the cache's
            // table entry, or converted from the mask form (on the host
            // pool); then the array row by row
            std::vector<const std::vector<uint32_t> *> wv((size_t)nj);
            std::vector<std::vector<uint32_t>> conv((size_t)nj);
            parallel_for(nj, [&](int b) {
                const size_t k = m.cells[(size_t)b];
                const std::vector<uint32_t> *w = &polys[(size_t)m.g][k];
                if (w->empty() && cache)
                    if (const auto *tw = cache->table_words(ref.slot, cells[(size_t)m.g][k].start))
                        w = tw;
                if (w->empty()) {
                    conv[(size_t)b] = words_from_mask16(m16[(size_t)m.g][k], (ref.mexp + 31) / 32);
                    w = &conv[(size_t)b];
                }
                wv[(size_t)b] = w;
            });
            parallel_for((wc + 15) / 16, [&](int blk) {
                for (int x = 16 * blk; x < std::min(wc, 16 * blk + 16); x++) {
                    uint32_t *row = cf + (size_t)x * (size_t)njp;
                    for (int b = 0; b < nj; b++)
                        row[b] = (size_t)x < wv[(size_t)b]->size() ? (*wv[(size_t)b])[(size_t)x] : 0u;
                }
            });
            gr.nj = nj;
            gr.njp = njp;
            gr.kpad = kpad;
            gr.pw = pw;
            gr.param = ref.slot;
            gr.n = ref.n;
            const int gi = (int)mgroups.size();
            mgroups.push_back(gr);
            mgroup_gen.push_back(m.g);
            for (int r0 = 0; r0 < nj; r0 += 32 * kMjRowTiles) {
                const int rt = std::min(kMjRowTiles, (nj - r0 + 31) / 32);
                for (int j0 = 0; j0 < ref.n; j0 += mj_words(rt))
                    mit[rt][m.phase].push_back(MJumpItem{gi, r0, j0, 0});
            }
            mjumps += nj;
        }
        int n0 = 0;
        for (const MGroup &m : mg)
            n0 += m.phase == 0;
        mrange[0].g0 = 0, mrange[0].g1 = n0;
        mrange[1].g0 = n0, mrange[1].g1 = (int)mg.size();
        mrange[2].g0 = 0, mrange[2].g1 = (int)mg.size();
        for (int rt = 1; rt <= kMjRowTiles; rt++) {
            mrange[2].i0[rt] = (int)mitems.size();
            for (int ph = 0; ph < 2; ph++) {
                mrange[ph].i0[rt] = (int)mitems.size();
                mitems.insert(mitems.end(), mit[rt][ph].begin(), mit[rt][ph].end());
                mrange[ph].i1[rt] = (int)mitems.size();
            }
            mrange[2].i1[rt] = (int)mitems.size();
        }
    }
    // Items, one per cell.
    std::vector<Item> list;
    std::vector<int> list_gen;
    std::vector<int> group;
    std::vector<uint64_t> arena_slot;  // per item: arena offset + 1 (0 = none)
    std::vector<uint64_t> cost;
    segs.clear();
    commits.clear();
    numbers = 0;
    jumps = 0;
    for (int g = 0; g < G; g++) {
        const GenRef &gr = gens[(size_t)g];
        const auto &cv = cells[(size_t)g];
        for (size_t k = 0; k < cv.size(); k++) {
            const Cell &cell = cv[k];
            Item it{};
            it.state_in = (uint64_t)(uintptr_t)gr.state_in;
            it.state_out = (k + 1 == cv.size()) ? (uint64_t)(uintptr_t)gr.state_out : 0;
            it.param = gr.slot;
            it.poly = UINT64_MAX;
            uint64_t n = 0;
            for (const Piece &pc : cell.pieces)
                n += pc.count;
            auto seg_of = [&](const Piece &pc) {
                Segment sg{};
                sg.count = pc.count;
                if (pc.req >= 0) {
                    const auto &r = reqs[(size_t)pc.req];
                    sg.out = (uint64_t)(uintptr_t)r.out + pc.out_off * (uint64_t)elem_bytes(gr.bits);
                    sg.kind = r.kind;
                }
                return sg;
            };
            if (cell.pieces.size() == 1 && cell.pieces[0].req >= 0) {
                const Segment sg = seg_of(cell.pieces[0]);
                it.out = sg.out;
                it.count = sg.count;
                it.kind = sg.kind;
            } else if (!cell.pieces.empty()) {
                it.segs = segs.size();  // index, fixed up below
                it.nsegs = (int32_t)cell.pieces.size();
                for (const Piece &pc : cell.pieces)
                    segs.push_back(seg_of(pc));
            }
            uint64_t slot = 0;
            if (cell.start > 0) {
                jumps++;
                if (jump_mode == 1) {
                    const auto &sp = slot_of.at({g, k});
                    slot = std::get<0>(sp) + 1;
                    it.parts = std::get<1>(sp);
                    it.part_bytes = std::get<2>(sp);
                } else {
                    const auto &pw = polys[(size_t)g][k];
                    it.poly = poly_host.size();  // word offset, fixed up below
                    it.poly_words = (int32_t)pw.size();
                    poly_host.insert(poly_host.end(), pw.begin(), pw.end());
                }
            }
            numbers += n;
            list.push_back(it);
            list_gen.push_back(g);
            arena_slot.push_back(slot);
            cost.push_back(n + (it.poly != UINT64_MAX ? 200000u : 0u));
            group.push_back(2 * group_of(gr) + phase_of(g, k));
            if (it.state_out)
                commits.push_back(g);
        }
    }
    // By (word size, class) group, phase 0 before phase 1, longest first.
    std::vector<int> order(list.size());
    for (size_t i = 0; i < order.size(); i++)
        order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (group[(size_t)a] != group[(size_t)b])
            return group[(size_t)a] < group[(size_t)b];
        return cost[(size_t)a] > cost[(size_t)b];
    });
    {
        std::vector<int> jorder(jitems.size());
        for (size_t i = 0; i < jorder.size(); i++)
            jorder[i] = (int)i;
        std::stable_sort(jorder.begin(), jorder.end(), [&](int a, int b) { return jgroup[(size_t)a] < jgroup[(size_t)b]; });
        std::vector<JumpItem> sorted;
        std::vector<int> sg;
        jitem_gen.clear();
        for (int i : jorder) {
            sorted.push_back(jitems[(size_t)i]);
            sg.push_back(jgroup[(size_t)i]);
            jitem_gen.push_back(jgen_list[(size_t)i]);
        }
        jitems.swap(sorted);
        jgroup.swap(sg);
    }
    // Workspace layout and device addresses.
    const size_t o_items = 0;
    const size_t o_segs = align256(o_items + list.size() * sizeof(Item));
    const size_t o_polys = align256(o_segs + segs.size() * sizeof(Segment));
    const size_t o_masks = align256(o_polys + poly_host.size() * 4);
    const size_t o_jitems = align256(o_masks + mask_host.size() * 8);
    const size_t o_mgroups = align256(o_jitems + jitems.size() * sizeof(JumpItem));
    const size_t o_mitems = align256(o_mgroups + mgroups.size() * sizeof(MJumpGroup));
    const size_t o_coef = align256(o_mitems + mitems.size() * sizeof(MJumpItem));
    const size_t work_bytes = align256(o_coef + coef_host.size() * 4);
    this->o_items = o_items;
    this->o_jitems = o_jitems;
    this->o_mgroups = o_mgroups;
    wait_upload();
    grow_device(d_work, work_cap, work_bytes, "hipMalloc plan workspace");
    grow_pinned(h_stage, stage_cap, work_bytes);
    if (arena_bytes)
        grow_device(d_jstates, jstates_cap, arena_bytes, "hipMalloc state arena");
    if (!d_queue) {
        hip_check(hipMalloc((void **)&d_queue, 64 * sizeof(uint32_t)), "hipMalloc queue");
        hip_check(hipMemset(d_queue, 0, 64 * sizeof(uint32_t)), "hipMemset queue");
    }
    const uint64_t dw = (uint64_t)(uintptr_t)d_work;
    d_items = reinterpret_cast<Item *>(d_work + o_items);
    d_jitems = reinterpret_cast<JumpItem *>(d_work + o_jitems);
    d_mgroups = reinterpret_cast<MJumpGroup *>(d_work + o_mgroups);
    d_mitems = reinterpret_cast<MJumpItem *>(d_work + o_mitems);
    items.clear();
    item_gen.clear();
    item_io.clear();
    for (int i : order) {
        Item it = list[(size_t)i];
        it.poly = it.poly == UINT64_MAX ? 0 : dw + o_polys + 4 * it.poly;
        if (it.nsegs > 0)
            it.segs = dw + o_segs + sizeof(Segment) * it.segs;
        if (arena_slot[(size_t)i])
            it.state_in = (uint64_t)(uintptr_t)d_jstates + arena_slot[(size_t)i] - 1;
        items.push_back(it);
        item_gen.push_back(list_gen[(size_t)i]);
        item_io.push_back((uint8_t)((arena_slot[(size_t)i] ? 0 : 1) | (it.state_out ? 2 : 0)));
    }
    all_commits = commits;
    for (auto &ji : jitems) {
        ji.masks = dw + o_masks + ji.masks;
        ji.states_out = (uint64_t)(uintptr_t)d_jstates + ji.states_out;
    }
    for (auto &gr : mgroups) {
        gr.planes += (uint64_t)(uintptr_t)d_jstates;
        gr.out += (uint64_t)(uintptr_t)d_jstates;
        gr.coef += dw + o_coef;
    }
    if (!items.empty())
        memcpy(h_stage + o_items, items.data(), items.size() * sizeof(Item));
    if (!segs.empty())
        memcpy(h_stage + o_segs, segs.data(), segs.size() * sizeof(Segment));
    if (!poly_host.empty())
        memcpy(h_stage + o_polys, poly_host.data(), poly_host.size() * 4);
    if (!mask_host.empty())
        memcpy(h_stage + o_masks, mask_host.data(), mask_host.size() * 8);
    if (!jitems.empty())
        memcpy(h_stage + o_jitems, jitems.data(), jitems.size() * sizeof(JumpItem));
    if (!mgroups.empty()) {
        memcpy(h_stage + o_mgroups, mgroups.data(), mgroups.size() * sizeof(MJumpGroup));
        memcpy(h_stage + o_mitems, mitems.data(), mitems.size() * sizeof(MJumpItem));
        memcpy(h_stage + o_coef, coef_host.data(), coef_host.size() * 4);
    }
    hip_check(hipMemcpyAsync(d_work, h_stage, work_bytes, hipMemcpyHostToDevice, opt.stream), "plan upload");
    if (!ev_upload)
        hip_check(hipEventCreateWithFlags(&ev_upload, hipEventDisableTiming), "hipEventCreateWithFlags");
    hip_check(hipEventRecord(ev_upload, opt.stream), "hipEventRecord");
    if (opt.sync)
        hip_check(hipStreamSynchronize(opt.stream), "plan upload sync");
    // Launches: one jump launch and one generation launch per (word size,
    // class) group and phase, and the same merged over the two phases (one
    // contiguous range per group: the serial order).
    auto group_launches = [&](const std::vector<int> &grp, bool jump) {
        std::vector<Launch> out;
        for (size_t i = 0; i < grp.size();) {
            const int gi = grp[i];
            size_t j = i;
            while (j < grp.size() && grp[j] == gi)
                j++;
            Launch L;
            L.phase = gi % 2;
            L.bits = gi / 2 >= 3 ? 64 : 32;
            L.cls = gi / 2 >= 3 ? gi / 2 - 3 : gi / 2;
            L.first = (int)i;
            L.count = (int)(j - i);
            L.layouts_ok = group_ok[gi / 2];
            if (jump) {
                L.layout = 1;
                L.grid = std::min(L.count, c.cus * c.jump_blocks_per_cu[L.bits == 64][L.cls]);
                L.dyn_lds = 0;
            } else {
                L.layout = c.layout_for(L.bits, L.cls, L.layouts_ok, one_round);
                const int cap = c.blocks_per_cu[L.bits == 64][L.cls][L.layout];
                // (the knob counts waves there; the occupancy cap counts
                // workgroups)
                const int knob = c.bpc[L.bits == 64][L.cls];
                int bpc = std::min(waves_scaled ? std::max(1, knob / layout_waves(L.layout)) : knob, cap);
                // Up to half a round more cells than slots (e.g. one cell per
                // generator for thousands of generators) run as one round
                // when the occupancy limit allows: a partial second round
                // would leave most waves idle.
                if (L.count > c.cus * bpc && L.count <= c.cus * cap && 2 * L.count <= 3 * c.cus * bpc)
                    bpc = (L.count + c.cus - 1) / c.cus;
                L.grid = std::min(L.count, c.cus * bpc);
                L.dyn_lds = c.lds_reserve(kernel_fn(L.bits, L.cls, L.layout), kernel_threads(L.bits, L.cls, L.layout), bpc);
            }
            out.push_back(L);
            i = j;
        }
        return out;
    };
    std::vector<int> sorted_group;
    for (int i : order)
        sorted_group.push_back(group[(size_t)i]);
    launches = group_launches(sorted_group, false);
    jlaunches = group_launches(jgroup, true);
    auto no_phase = [](std::vector<int> v) {
        for (int &x : v)
            x &= ~1;
        return v;
    };
    mlaunches = group_launches(no_phase(sorted_group), false);
    mjlaunches = group_launches(no_phase(jgroup), true);
    if (2 * (launches.size() + jlaunches.size() + mlaunches.size() + mjlaunches.size()) > 64)
        throw Error("too many launches in one plan");
    has_phase1 = false;
    for (const Launch &L : launches)
        has_phase1 |= L.phase == 1;
    overlap = opt.overlap >= 0 ? opt.overlap != 0 : c.overlap;
    segment_used = S;
    plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void Plan::wait_upload() {
    if (ev_upload)
        hip_check(hipEventSynchronize(ev_upload), "hipEventSynchronize");
}

void Plan::retarget(const std::vector<GenRef> &gens, hipStream_t s) {
    if (!d_work)
        throw Error("plan has no work list");
    for (size_t i = 0; i < items.size(); i++) {
        const GenRef &gr = gens.at((size_t)item_gen[i]);
        if (item_io[i] & 1)
            items[i].state_in = (uint64_t)(uintptr_t)gr.state_in;
        if (item_io[i] & 2)
            items[i].state_out = (uint64_t)(uintptr_t)gr.state_out;
    }
    for (size_t i = 0; i < jitems.size(); i++)
        jitems[i].state_in = (uint64_t)(uintptr_t)gens.at((size_t)jitem_gen[i]).state_in;
    for (size_t i = 0; i < mgroups.size(); i++)
        mgroups[i].state_in = (uint64_t)(uintptr_t)gens.at((size_t)mgroup_gen[i]).state_in;
    wait_upload();
    const size_t mb = mgroups.size() * sizeof(MJumpGroup);
    if (mb) {
        memcpy(h_stage + o_mgroups, mgroups.data(), mb);
        hip_check(hipMemcpyAsync(d_work + o_mgroups, h_stage + o_mgroups, mb, hipMemcpyHostToDevice, s), "plan upload");
    }
    const size_t ib = items.size() * sizeof(Item), jb = jitems.size() * sizeof(JumpItem);
    if (ib) {
        memcpy(h_stage + o_items, items.data(), ib);
        hip_check(hipMemcpyAsync(d_work + o_items, h_stage + o_items, ib, hipMemcpyHostToDevice, s), "plan upload");
    }
    if (jb) {
        memcpy(h_stage + o_jitems, jitems.data(), jb);
        hip_check(hipMemcpyAsync(d_work + o_jitems, h_stage + o_jitems, jb, hipMemcpyHostToDevice, s), "plan upload");
    }
    if (!ev_upload)
        hip_check(hipEventCreateWithFlags(&ev_upload, hipEventDisableTiming), "hipEventCreateWithFlags");
    hip_check(hipEventRecord(ev_upload, s), "hipEventRecord");
    commits = all_commits;
}

void Plan::reserve(size_t work_bytes, size_t jstates_bytes) {
    grow_device(d_work, work_cap, work_bytes, "hipMalloc plan workspace");
    grow_pinned(h_stage, stage_cap, work_bytes);
    grow_device(d_jstates, jstates_cap, jstates_bytes, "hipMalloc state arena");
    if (!d_queue) {
        hip_check(hipMalloc((void **)&d_queue, 64 * sizeof(uint32_t)), "hipMalloc queue");
        hip_check(hipMemset(d_queue, 0, 64 * sizeof(uint32_t)), "hipMemset queue");
    }
}

void Plan::launch(hipStream_t s) {
    if (!two_phase()) {
        launch_part(0, s);
        launch_part(1, s);
        return;
    }
    ensure_side();
    launch_phase(0, 0, s);
    hip_check(hipEventRecord(ev_fork, s), "hipEventRecord");
    hip_check(hipStreamWaitEvent(side, ev_fork, 0), "hipStreamWaitEvent");
    launch_phase(0, 1, side);
    launch_phase(1, 0, s);
    launch_phase(1, 1, side);
    hip_check(hipEventRecord(ev_join, side), "hipEventRecord");
    hip_check(hipStreamWaitEvent(s, ev_join, 0), "hipStreamWaitEvent");
}

void Plan::ensure_side() {
    if (side)
        return;
    hip_check(hipStreamCreateWithFlags(&side, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    hip_check(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming), "hipEventCreateWithFlags");
    hip_check(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming), "hipEventCreateWithFlags");
}

void Plan::launch_part(int part, hipStream_t s) { run_launches(part, -1, s); }

void Plan::launch_phase(int part, int phase, hipStream_t s) { run_launches(part, phase, s); }

// Work-queue counter pairs: [launches | jlaunches | mlaunches | mjlaunches].
void Plan::run_launches(int part, int phase, hipStream_t s) {
    const bool merged = phase < 0;
    const auto &J = merged ? mjlaunches : jlaunches;
    const auto &I = merged ? mlaunches : launches;
    const size_t iq = merged ? 2 * (launches.size() + jlaunches.size()) : 0;
    const size_t jq = iq + 2 * I.size();
    if (part == 0 && !mgroups.empty()) {
        // matrix-core jumps: the planes of the phase's groups, then the
        // products by row-tile count
        const MRange &R = mrange[merged ? 2 : phase];
        if (R.g1 > R.g0) {
            hip_check(launch_planes32(d_mgroups + R.g0, R.g1 - R.g0, ctx->d_params, s), "planes kernel launch");
            for (int rt = 1; rt <= kMjRowTiles; rt++)
                hip_check(launch_jump_mfma32(d_mgroups, d_mitems + R.i0[rt], R.i1[rt] - R.i0[rt], rt, s),
                          "matrix-core jump launch");
        }
    }
    for (size_t k = 0; part == 0 && k < J.size(); k++) {
        const Launch &L = J[k];
        if (!merged && L.phase != phase)
            continue;
        hip_check(launch_jumps(L.bits, L.cls, d_jitems + L.first, L.count, ctx->d_params, d_queue + jq + 2 * k, L.grid,
                               L.dyn_lds, s),
                  "jump kernel launch");
    }
    for (size_t k = 0; part == 1 && k < I.size(); k++) {
        const Launch &L = I[k];
        if (!merged && L.phase != phase)
            continue;
        hip_check(launch_items(L.bits, L.cls, L.layout, d_items + L.first, L.count, ctx->d_params, d_queue + iq + 2 * k,
                               L.grid, L.dyn_lds, s),
                  "kernel launch");
    }
}

void Plan::configure(int layout, int bpc) {
    drop_graph();
    std::shared_lock<std::shared_mutex> tune_lock(ctx->tune_mu);
    for (auto *Ls : {&launches, &mlaunches})
    for (auto &L : *Ls) {
        if (layout >= 0) {
            if (layout >= kLayouts)
                throw Error("unknown layout");
            // multi-wave cells only where the launch admits them
            L.layout = Context::layout_within(layout, L.layouts_ok);
        }
        int cap = ctx->blocks_per_cu[L.bits == 64][L.cls][L.layout];
        cap = std::min(cap, bpc > 0 ? bpc
                                    : waves_scaled ? std::max(1, ctx->bpc[L.bits == 64][L.cls] / layout_waves(L.layout))
                                                   : ctx->bpc[L.bits == 64][L.cls]);
        L.grid = std::min(L.count, ctx->cus * cap);
        L.dyn_lds = ctx->lds_reserve(kernel_fn(L.bits, L.cls, L.layout), kernel_threads(L.bits, L.cls, L.layout), cap);
    }
}

void Plan::drop_graph() {
    if (graph_exec)
        (void)hipGraphExecDestroy(graph_exec);
    if (graph)
        (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
}

void Plan::capture(hipStream_t s) {
    if (!s)
        throw Error("capture needs a stream other than the null stream");
    drop_graph();
    if (two_phase())
        ensure_side();  // created outside the capture
    hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
        launch(s);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g)
            (void)hipGraphDestroy(g);
        throw;
    }
    hip_check(hipStreamEndCapture(s, &graph), "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
}

void Plan::launch_graph(hipStream_t s) {
    if (!graph_exec)
        throw Error("plan has no captured graph");
    hip_check(hipGraphLaunch(graph_exec, s), "hipGraphLaunch");
}

float Plan::autotune(hipStream_t s, int reps, int *best_layout, int *best_bpc) {
    if (launches.empty())
        return 0.f;
    reps = std::max(1, reps);
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    float best = -1.f;
    int bl = -1, bb = -1;
    try {
        launch_part(0, s);  // start states stay valid for every trial
        for (int layout = 0; layout < kLayouts; layout++) {
            bool any = layout < 2;
            for (const auto &L : launches)
                any = any || (L.layouts_ok >> layout & 1);
            if (!any)
                continue;
            int last_cap = -1;
            for (int bpc : {4, 6, 8, 12, 16, 20, 24, 32}) {
                int cap = bpc;
                for (const auto &L : launches)
                    cap = std::min(cap, ctx->blocks_per_cu[L.bits == 64][L.cls][layout]);
                if (cap == last_cap)
                    continue;
                last_cap = cap;
                configure(layout, cap);
                launch_part(1, s);  // warm-up
                hip_check(hipEventRecord(e0, s), "hipEventRecord");
                for (int r = 0; r < reps; r++)
                    launch_part(1, s);
                hip_check(hipEventRecord(e1, s), "hipEventRecord");
                hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                float ms = 0.f;
                hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                ms /= (float)reps;
                if (best < 0.f || ms < best) {
                    best = ms;
                    bl = layout;
                    bb = cap;
                }
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    configure(bl, bb);
    std::unique_lock<std::shared_mutex> tune_lock(ctx->tune_mu);
    ctx->tune_epoch++;
    for (const auto &L : launches) {
        if (L.layout == bl)
            (one_round_grid ? ctx->layout_one_round : ctx->layout)[L.bits == 64][L.cls] = bl;
        ctx->bpc[L.bits == 64][L.cls] = bb;
    }
    if (best_layout)
        *best_layout = bl;
    if (best_bpc)
        *best_bpc = bb;
    return best;
}

void Plan::free_device() {
    drop_graph();
    if (ev_upload)
        (void)hipEventDestroy(ev_upload);
    ev_upload = nullptr;
    if (ev_fork)
        (void)hipEventDestroy(ev_fork);
    if (ev_join)
        (void)hipEventDestroy(ev_join);
    if (side)
        (void)hipStreamDestroy(side);
    ev_fork = ev_join = nullptr;
    side = nullptr;
    if (d_work)
        (void)hipFree(d_work);
    if (d_jstates)
        (void)hipFree(d_jstates);
    if (d_queue)
        (void)hipFree(d_queue);
    if (h_stage)
        (void)hipHostFree(h_stage);
    d_work = d_jstates = h_stage = nullptr;
    d_queue = nullptr;
    work_cap = jstates_cap = stage_cap = 0;
    d_items = nullptr;
    d_jitems = nullptr;
    d_mgroups = nullptr;
    d_mitems = nullptr;
}

Plan::~Plan() { free_device(); }

}  // namespace mtgp

using namespace mtgp;

// ------------------------------------------------------------------ C API

#define API_BEGIN try {
#define API_END(fail)                   \
    }                                   \
    catch (const std::exception &e) {   \
        set_error(e.what());            \
        return fail;                    \
    }

// A plan over a genset remembers, per committed generator, the slot its
// launch writes and the generator's epoch when the plan was built or
// rebased; commit refuses a plan that was never launched or whose
// generators' states have changed since (which would rewind them).
struct mtgp_plan {
    mtgp_genset *set;
    Plan plan;
    std::vector<int> out_slot;
    std::vector<uint64_t> epochs;
    bool launched = false;
    void record_targets() {
        out_slot.clear();
        epochs.clear();
        for (int g : plan.commits) {
            out_slot.push_back(set->cur[(size_t)g] ^ 1);
            epochs.push_back(set->epoch[(size_t)g]);
        }
        launched = false;
    }
};

// Applies a jump polynomial (hex) to device window `in`, result in `out`.
// A single jump on one wave is latency-bound (the fold of every chunk of
// the polynomial in sequence), so it is split into partial jumps over chunk
// ranges, one wave each, whose windows a second launch XORs together (the
// identity polynomial with JumpItem.in_parts) into `out`.
// One jump of a window by a jump string: the coefficients split over up to
// jump_max_parts waves by chunk range, their partial windows XORed by a
// second launch.  In the context's workspace [masks | items | queues |
// window in | partial windows | window out]; everything the kernels read
// goes up in one copy from the pinned staging copy of that prefix.  With
// host_io, `in` and `out` are host windows staged through the workspace
// (one copy each way), else device windows.
static void device_jump(Context &c, int slot, int bits, int mexp, const void *in, void *out, const char *hex,
                        bool host_io) {
    gf2::Poly P;
    if (!gf2::from_hex(hex, P))
        throw Error("jump string contains a non-hex character");
    // Applied literally, as mtgp32_fast_jump does: every coefficient of the
    // string, whatever its degree (the kernel runs the sequence that far).
    const std::vector<uint32_t> words = gf2::to_words32(P, std::max<int>(1, (int)((strlen(hex) * 4 + 31) / 32)));
    std::vector<uint64_t> masks;
    const int chunks = append_jump_masks({&words}, masks);
    const size_t id_off = masks.size();
    const std::vector<uint32_t> one(1, 1u);
    append_jump_masks({&one}, masks);
    const int parts = std::max(1, std::min(chunks, c.jump_max_parts));
    const int wbytes = state_words(bits, mexp) * (bits / 8);
    const size_t stride = (wbytes + 255) & ~(size_t)255;
    const size_t o_items = (masks.size() * 8 + 255) & ~(size_t)255;
    const size_t o_queue = o_items + (((size_t)(parts + 1) * sizeof(JumpItem) + 255) & ~(size_t)255);
    const size_t o_in = o_queue + 256;
    const size_t o_parts = o_in + stride;
    const size_t o_out = o_parts + (size_t)parts * stride;
    const size_t up = host_io ? o_in + (size_t)wbytes : o_in;  // bytes uploaded
    std::lock_guard<std::mutex> lk(c.jump_mu);
    grow_device(c.jump_d, c.jump_dcap, o_out + stride, "hipMalloc jump workspace");
    grow_pinned(c.jump_h, c.jump_hcap, o_out + stride);
    char *d = c.jump_d, *h = c.jump_h;
    const char *win_in = host_io ? d + o_in : static_cast<const char *>(in);
    char *win_out = host_io ? d + o_out : static_cast<char *>(out);
    memset(h, 0, o_in);
    memcpy(h, masks.data(), masks.size() * 8);
    JumpItem *items = reinterpret_cast<JumpItem *>(h + o_items);
    for (int q = 0; q < parts; q++) {
        JumpItem &ji = items[q];
        ji.state_in = (uint64_t)(uintptr_t)win_in;
        ji.masks = (uint64_t)(uintptr_t)d;
        ji.states_out = (uint64_t)(uintptr_t)(d + o_parts + (size_t)q * stride);
        ji.nb = 1;
        ji.c0 = (int32_t)((long)chunks * q / parts);
        ji.c1 = (int32_t)((long)chunks * (q + 1) / parts);
        ji.param = slot;
        ji.state_bytes = wbytes;
    }
    JumpItem &cj = items[parts];
    cj.state_in = (uint64_t)(uintptr_t)(d + o_parts);
    cj.in_parts = parts;
    cj.in_part_bytes = stride;
    cj.masks = (uint64_t)(uintptr_t)(d + id_off * 8);
    cj.states_out = (uint64_t)(uintptr_t)win_out;
    cj.nb = 1;
    cj.c0 = 0;
    cj.c1 = 1;
    cj.param = slot;
    cj.state_bytes = wbytes;
    if (host_io)
        memcpy(h + o_in, in, (size_t)wbytes);
    // the queue pairs go up as zeros with the rest
    hip_check(hipMemcpyAsync(d, h, up, hipMemcpyHostToDevice, nullptr), "hipMemcpyAsync jump workspace");
    const int cls = c.class_of(slot);
    const JumpItem *d_items = reinterpret_cast<const JumpItem *>(d + o_items);
    uint32_t *d_q = reinterpret_cast<uint32_t *>(d + o_queue);
    hip_check(launch_jumps(bits, cls, d_items, parts, c.d_params, d_q, parts, 0, nullptr), "jump kernel launch");
    hip_check(launch_jumps(bits, cls, d_items + parts, 1, c.d_params, d_q + 2, 1, 0, nullptr), "jump kernel launch");
    if (host_io)
        hip_check(hipMemcpyAsync(h + o_out, d + o_out, (size_t)wbytes, hipMemcpyDeviceToHost, nullptr),
                  "hipMemcpyAsync jumped window");
    hip_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    if (host_io)
        memcpy(out, h + o_out, (size_t)wbytes);
}

template <class Gen, class Word>
static int fill_impl(Gen *g, int kind, void *out, size_t n, int bits) {
    Context &c = context();
    if (kind < 0 || kind > 4)
        throw Error("unknown output kind");
    if (!g || !g->status)
        throw Error("generator is not initialised");
    const int N = g->status->size;
    DevParam rec = dev_param_of(g->params);
    const int slot = c.slot_of(rec, g->params.poly_sha1);
    std::vector<Word> w((size_t)N);
    for (int j = 0; j < N; j++)
        w[(size_t)j] = g->status->array[(g->status->idx - N + 1 + j + g->status->large_size) & g->status->large_mask];
    const size_t eb = (size_t)bits / 8;
    bool dev_out = false;
    if (out) {
        hipPointerAttribute_t attr;
        if (hipPointerGetAttributes(&attr, out) == hipSuccess &&
            (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged))
            dev_out = true;
        (void)hipGetLastError();
    }
    if (n && !out)
        throw Error("NULL output buffer");
    std::lock_guard<std::mutex> fill_lock(c.fill_mu);
    if (!c.fill_plan)
        c.fill_plan = new Plan();
    Plan &plan = *c.fill_plan;
    grow_device(c.fill_states, c.fill_states_cap, 2 * (size_t)N * sizeof(Word), "hipMalloc fill states");
    char *d_states = c.fill_states;
    void *d_out = out;
    if (n && !dev_out) {
        grow_device(c.fill_out, c.fill_out_cap, n * eb, "hipMalloc fill output");
        d_out = c.fill_out;
    }
    // Windows (and host outputs up to kFillStage bytes) go through pinned
    // staging: one asynchronous copy each way and one synchronisation.
    constexpr size_t kFillStage = (size_t)64 << 20;
    const size_t wb = (size_t)N * sizeof(Word);
    const bool stage_out = n && !dev_out && n * eb <= kFillStage;
    grow_pinned(c.fill_h, c.fill_hcap, 2 * wb + (stage_out ? n * eb : 0));
    char *h_in = c.fill_h, *h_win = c.fill_h + wb, *h_out = c.fill_h + 2 * wb;
    memcpy(h_in, w.data(), wb);
    hip_check(hipMemcpyAsync(d_states, h_in, wb, hipMemcpyHostToDevice, nullptr), "hipMemcpyAsync window");
    GenRef r;
    r.bits = bits;
    r.mexp = g->params.mexp;
    r.n = N;
    r.slot = slot;
    r.state_in = d_states;
    r.state_out = d_states + (size_t)N * sizeof(Word);
    Context::FillShape shape;
    shape.slot = slot;
    shape.bits = bits;
    shape.kind = kind;
    shape.n = n;
    shape.out = d_out;
    shape.states = d_states;
    {
        std::shared_lock<std::shared_mutex> tune_lock(c.tune_mu);
        shape.epoch = c.tune_epoch;
    }
    if (!(shape == c.fill_shape)) {
        c.fill_shape = Context::FillShape();  // a failed build leaves no reusable shape
        mtgp_request_t q{0, kind, (uint64_t)n, 0, d_out};
        plan.build(c, {r}, {q}, 0);
        c.fill_shape = shape;
    }
    plan.launch(0);
    hip_check(hipMemcpyAsync(h_win, r.state_out, wb, hipMemcpyDeviceToHost, nullptr), "hipMemcpyAsync window");
    if (stage_out)
        hip_check(hipMemcpyAsync(h_out, d_out, n * eb, hipMemcpyDeviceToHost, nullptr), "hipMemcpyAsync output");
    hip_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    if (stage_out)
        memcpy(out, h_out, n * eb);
    else if (n && !dev_out)
        hip_check(hipMemcpy(out, d_out, n * eb, hipMemcpyDeviceToHost), "hipMemcpy");
    memcpy(g->status->array, h_win, wb);
    g->status->idx = N - 1;
    return 0;
}


template <class Gen, class Word>
static void fast_jump_impl(Gen *g, const char *hex, int bits) {
    Context &c = context();
    const int N = g->status->size;
    DevParam rec = dev_param_of(g->params);
    const int slot = c.slot_of(rec, g->params.poly_sha1);
    std::vector<Word> w((size_t)N);
    for (int j = 0; j < N; j++)
        w[(size_t)j] = g->status->array[(g->status->idx - N + 1 + j + g->status->large_size) & g->status->large_mask];
    device_jump(c, slot, bits, g->params.mexp, w.data(), w.data(), hex, true);
    memcpy(g->status->array, w.data(), (size_t)N * sizeof(Word));
    g->status->idx = N - 1;
}


static int put_string(const std::string &s, char *buf, size_t cap) {
    if (!buf || s.size() + 1 > cap) {
        set_error("buffer too small");
        return -1;
    }
    memcpy(buf, s.c_str(), s.size() + 1);
    return (int)s.size();
}

// Characteristic and jump polynomials are host arithmetic: no GPU needed.
template <class Params>
static int charpoly_impl(const Params &p, char *buf, size_t cap) {
    auto m = modulus_for(dev_param_of(p), p.poly_sha1);
    return put_string(gf2::to_hex(m->poly()), buf, cap);
}

template <class Params>
static int jump_impl(const Params &p, const uint64_t *steps, int nwords, char *buf, size_t cap) {
    if (nwords <= 0)
        throw Error("empty step count");
    auto m = modulus_for(dev_param_of(p), p.poly_sha1);
    std::vector<uint64_t> e(steps, steps + nwords);
    return put_string(gf2::to_hex(m->xpow(e)), buf, cap);
}

extern "C" {

const char *mtgp_last_error(void) { return last_error(); }

int mtgp_set_device(int device) {
    API_BEGIN
    int count = 0;
    hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (device < 0 || device >= count)
        throw Error("no such device");
    select_device(device);
    context();
    return 0;
    API_END(-1)
}

int mtgp_device_info(int *cus, size_t *hbm) {
    API_BEGIN
    Context &c = context();
    if (cus)
        *cus = c.cus;
    if (hbm)
        *hbm = c.hbm_bytes;
    return 0;
    API_END(-1)
}

namespace {
// The planner's tuning knobs by name (each also read from an environment
// variable at start-up).
double *tuning_knob(Context &c, const std::string &name) {
    if (name == "segment_split")
        return &c.segment_split;
    if (name == "segment_rounds")
        return &c.segment_rounds;
    if (name == "small_rounds")
        return &c.small_rounds;
    if (name == "two_size_rho")
        return &c.two_size_rho;
    if (name == "jump_floor_ms")
        return &c.jump_floor_ms;
    return nullptr;
}
}  // namespace

int mtgp_tuning_set(const char *name, double value) {
    API_BEGIN
    Context &c = context();
    const std::string n = name ? name : "";
    std::unique_lock<std::shared_mutex> lk(c.tune_mu);
    c.tune_epoch++;
    if (n == "overlap") {
        c.overlap = value != 0;
        return 0;
    }
    if (n == "jump_waves_per_cu") {
        if (value < 1 || value > 64)
            throw Error("jump_waves_per_cu outside [1, 64]");
        c.jump_waves_per_cu = (int)value;
        return 0;
    }
    if (n == "jump_max_parts") {
        if (value < 1 || value > 64)
            throw Error("jump_max_parts outside [1, 64]");
        c.jump_max_parts = (int)value;
        return 0;
    }
    if (n == "jump_mfma_min") {
        if (!(value >= 0) || value > (1 << 30))
            throw Error("jump_mfma_min outside [0, 2^30]");
        c.jump_mfma_min = (int)value;
        return 0;
    }
    double *k = tuning_knob(c, n);
    if (!k)
        throw Error("unknown tuning knob '" + n + "'");
    if (!(value >= 0) || (n == "segment_split" && value >= 1) || (n == "segment_rounds" && value < 0.25) ||
        
        (n == "small_rounds" && (value < 0.1 || value > 8)))
        throw Error("value out of range for '" + n + "'");
    *k = value;
    return 0;
    API_END(-1)
}

int mtgp_tuning_get(const char *name, double *value) {
    API_BEGIN
    Context &c = context();
    const std::string n = name ? name : "";
    std::shared_lock<std::shared_mutex> lk(c.tune_mu);
    double v;
    if (n == "overlap")
        v = c.overlap ? 1 : 0;
    else if (n == "jump_waves_per_cu")
        v = c.jump_waves_per_cu;
    else if (n == "jump_max_parts")
        v = c.jump_max_parts;
    else if (n == "jump_mfma_min")
        v = c.jump_mfma_min;
    else if (double *k = tuning_knob(c, n))
        v = *k;
    else
        throw Error("unknown tuning knob '" + n + "'");
    if (value)
        *value = v;
    return 0;
    API_END(-1)
}

void *mtgp_dev_malloc(size_t bytes) {
    API_BEGIN
    context();
    void *p = nullptr;
    hip_check(hipMalloc(&p, bytes ? bytes : 1), "hipMalloc");
    return p;
    API_END(nullptr)
}

int mtgp_dev_free(void *p) {
    API_BEGIN
    hip_check(hipFree(p), "hipFree");
    return 0;
    API_END(-1)
}

int mtgp_memcpy_htod(void *dst, const void *src, size_t bytes) {
    API_BEGIN
    hip_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    return 0;
    API_END(-1)
}

int mtgp_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    API_BEGIN
    hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    return 0;
    API_END(-1)
}

int mtgp_device_copy(void *dst, const void *src, size_t bytes, void *stream) {
    API_BEGIN
    Context &c = context();
    if (((uintptr_t)dst | (uintptr_t)src | bytes) & 15)
        throw Error("mtgp_device_copy needs 16-byte aligned pointers and size");
    if (bytes)
        hip_check(launch_copy16(src, dst, bytes / 16, c.cus * 8, (hipStream_t)stream), "copy launch");
    return 0;
    API_END(-1)
}

int mtgp_device_fill(void *dst, uint32_t value, size_t bytes, void *stream) {
    API_BEGIN
    Context &c = context();
    if (((uintptr_t)dst | bytes) & 3)
        throw Error("mtgp_device_fill needs 4-byte aligned pointer and size");
    hip_check(launch_fill(dst, bytes / 4, value, c.cus * 16, (hipStream_t)stream), "fill launch");
    return 0;
    API_END(-1)
}

int mtgp_synchronize(void) {
    API_BEGIN
    context();
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return 0;
    API_END(-1)
}

void *mtgp_stream_create(void) {
    API_BEGIN
    context();
    hipStream_t s;
    hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    return (void *)s;
    API_END(nullptr)
}

int mtgp_stream_destroy(void *s) {
    API_BEGIN
    hip_check(hipStreamDestroy((hipStream_t)s), "hipStreamDestroy");
    return 0;
    API_END(-1)
}

int mtgp_stream_synchronize(void *s) {
    API_BEGIN
    hip_check(hipStreamSynchronize((hipStream_t)s), "hipStreamSynchronize");
    return 0;
    API_END(-1)
}

void *mtgp_event_create(void) {
    API_BEGIN
    context();
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return (void *)e;
    API_END(nullptr)
}

int mtgp_event_destroy(void *e) {
    API_BEGIN
    hip_check(hipEventDestroy((hipEvent_t)e), "hipEventDestroy");
    return 0;
    API_END(-1)
}

int mtgp_event_record(void *e, void *s) {
    API_BEGIN
    hip_check(hipEventRecord((hipEvent_t)e, (hipStream_t)s), "hipEventRecord");
    return 0;
    API_END(-1)
}

int mtgp_event_elapsed_ms(void *a, void *b, float *ms) {
    API_BEGIN
    hip_check(hipEventSynchronize((hipEvent_t)b), "hipEventSynchronize");
    hip_check(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b), "hipEventElapsedTime");
    return 0;
    API_END(-1)
}

// ------------------------------------------------------------ gensets

// known_slots: table slots when the records come from the tables.
static mtgp_genset_t *genset_of(int count, const int *bits, const std::vector<const void *> &recs,
                                const std::vector<int> *known_slots = nullptr) {
    if (count <= 0)
        throw Error("generator count must be positive");
    Context &c = context();
    std::unique_ptr<mtgp_genset> s(new mtgp_genset());
    s->ctx = &c;
    size_t bytes = 0;
    for (int g = 0; g < count; g++) {
        int slot, mexp;
        if (bits[g] == 32) {
            const auto &p = *static_cast<const mtgp32_params_fast_t *>(recs[(size_t)g]);
            slot = known_slots ? (*known_slots)[(size_t)g] : c.slot_of(dev_param_of(p), p.poly_sha1);
            mexp = p.mexp;
            s->p32.push_back(p);
            s->p64.emplace_back();
        } else if (bits[g] == 64) {
            const auto &p = *static_cast<const mtgp64_params_fast_t *>(recs[(size_t)g]);
            check64(p);
            slot = known_slots ? (*known_slots)[(size_t)g] : c.slot_of(dev_param_of(p), p.poly_sha1);
            mexp = p.mexp;
            s->p64.push_back(p);
            s->p32.emplace_back();
        } else {
            throw Error("word size must be 32 or 64");
        }
        if (c.class_of(slot) < 0)
            throw Error("parameter set fits no device kernel (MEXP " + std::to_string(mexp) + ")");
        s->bits.push_back(bits[g]);
        s->mexp.push_back(mexp);
        s->slot.push_back(slot);
        s->n.push_back(state_words(bits[g], mexp));
        s->cur.push_back(0);
        s->epoch.push_back(0);
        s->off.push_back(bytes);
        bytes += 2 * (size_t)s->n.back() * (size_t)(bits[g] / 8);
        bytes = (bytes + 255) & ~(size_t)255;
    }
    hip_check(hipMalloc((void **)&s->d_states, bytes), "hipMalloc states");
    hip_check(hipMemset(s->d_states, 0, bytes), "hipMemset states");
    return s.release();
}

mtgp_genset_t *mtgp_genset_create(int count, const int *bits, const int *mexp, const int *index) {
    API_BEGIN
    if (count <= 0)
        throw Error("generator count must be positive");
    std::vector<const void *> recs;
    std::vector<int> slots;
    for (int g = 0; g < count; g++) {
        const int t = table_of(bits[g], mexp[g]);
        if (t < 0 || index[g] < 0 || index[g] >= kSetsPerTable)
            throw Error("no parameter set mtgp" + std::to_string(bits[g]) + "-" + std::to_string(mexp[g]) + " #" +
                        std::to_string(index[g]));
        static mtgp32_params_fast_t *const t32[3] = {mtgp32_params_fast_11213, mtgp32_params_fast_23209,
                                                     mtgp32_params_fast_44497};
        static mtgp64_params_fast_t *const t64[3] = {mtgp64_params_fast_23209, mtgp64_params_fast_44497,
                                                     mtgp64_params_fast_110503};
        recs.push_back(t < 3 ? (const void *)&t32[t][index[g]] : (const void *)&t64[t - 3][index[g]]);
        slots.push_back(global_set_id(bits[g], mexp[g], index[g]));
    }
    return genset_of(count, bits, recs, &slots);
    API_END(nullptr)
}

mtgp_genset_t *mtgp_genset_create_params(int count, const int *bits, const void *const *params) {
    API_BEGIN
    if (count <= 0)
        throw Error("generator count must be positive");
    return genset_of(count, bits, std::vector<const void *>(params, params + count));
    API_END(nullptr)
}

void mtgp_genset_destroy(mtgp_genset_t *s) { delete s; }

int mtgp_genset_count(const mtgp_genset_t *s) { return (int)s->bits.size(); }

int mtgp_genset_state_words(const mtgp_genset_t *s, int g) {
    if (g < 0 || g >= (int)s->bits.size()) {
        set_error("generator index out of range");
        return -1;
    }
    return s->n[(size_t)g];
}

static void upload_window(mtgp_genset *s, int g, const void *w) {
    s->set_cur(g, 0);
    hip_check(hipMemcpy(s->state(g, 0), w, (size_t)s->n[(size_t)g] * (size_t)(s->bits[(size_t)g] / 8),
                        hipMemcpyHostToDevice),
              "hipMemcpy state");
}

int mtgp_genset_init_seed(mtgp_genset_t *s, const uint64_t *seeds) {
    API_BEGIN
    for (int g = 0; g < (int)s->bits.size(); g++) {
        if (s->bits[(size_t)g] == 32) {
            std::vector<uint32_t> w((size_t)s->n[(size_t)g]);
            mtgp32_init_state(w.data(), &s->p32[(size_t)g], (uint32_t)seeds[g]);
            upload_window(s, g, w.data());
        } else {
            std::vector<uint64_t> w((size_t)s->n[(size_t)g]);
            mtgp64_init_state(w.data(), &s->p64[(size_t)g], seeds[g]);
            upload_window(s, g, w.data());
        }
    }
    return 0;
    API_END(-1)
}

// Device initialisation (init_kernels.hip): mode 0 seeds, mode 1 keys.
static void device_init(mtgp_genset *s, const uint64_t *seeds, const uint64_t *keys, const int *offsets,
                        const int *lengths, hipStream_t st) {
    const int G = (int)s->bits.size();
    std::vector<InitItem> items((size_t)G);
    size_t nkeys = 0;
    if (keys)
        for (int g = 0; g < G; g++) {
            if (offsets[g] < 0 || lengths[g] < 0)
                throw Error("negative key offset or length");
            nkeys = std::max(nkeys, (size_t)offsets[g] + (size_t)lengths[g]);
        }
    uint64_t *d_keys = nullptr;
    InitItem *d_items = nullptr;
    try {
        hip_check(hipMalloc((void **)&d_items, items.size() * sizeof(InitItem)), "hipMalloc init items");
        if (nkeys) {
            hip_check(hipMalloc((void **)&d_keys, nkeys * 8), "hipMalloc keys");
            hip_check(hipMemcpy(d_keys, keys, nkeys * 8, hipMemcpyHostToDevice), "hipMemcpy keys");
        }
        for (int g = 0; g < G; g++) {
            InitItem &it = items[(size_t)g];
            it.state = (uint64_t)(uintptr_t)s->state(g, s->cur[(size_t)g]);
            it.n = s->n[(size_t)g];
            it.bits = s->bits[(size_t)g];
            if (it.bits == 32) {
                const auto &p = s->p32[(size_t)g];
                it.hidden = p.tbl[4] ^ (p.tbl[8] << 16);
            } else {
                const auto &p = s->p64[(size_t)g];
                it.hidden = p.tbl[4] ^ (p.tbl[8] << 16);
            }
            if (keys) {
                it.mode = 1;
                it.key_len = lengths[g];
                it.seed = (uint64_t)(uintptr_t)(d_keys + offsets[g]);
            } else {
                it.mode = 0;
                it.seed = it.bits == 32 ? (uint64_t)(uint32_t)seeds[g] : seeds[g];
            }
        }
        hip_check(hipMemcpy(d_items, items.data(), items.size() * sizeof(InitItem), hipMemcpyHostToDevice),
                  "hipMemcpy init items");
        hip_check(launch_init(d_items, G, st), "init kernel launch");
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    } catch (...) {
        if (d_items)
            (void)hipFree(d_items);
        if (d_keys)
            (void)hipFree(d_keys);
        throw;
    }
    (void)hipFree(d_items);
    if (d_keys)
        (void)hipFree(d_keys);
    for (int g = 0; g < G; g++)
        s->epoch[(size_t)g]++;
}

int mtgp_genset_init_seed_device(mtgp_genset_t *s, const uint64_t *seeds, void *stream) {
    API_BEGIN
    device_init(s, seeds, nullptr, nullptr, nullptr, (hipStream_t)stream);
    return 0;
    API_END(-1)
}

int mtgp_genset_init_array_device(mtgp_genset_t *s, const uint64_t *keys, const int *offsets, const int *lengths,
                                  void *stream) {
    API_BEGIN
    device_init(s, nullptr, keys, offsets, lengths, (hipStream_t)stream);
    return 0;
    API_END(-1)
}

int mtgp_genset_init_array(mtgp_genset_t *s, const uint64_t *keys, const int *offsets, const int *lengths) {
    API_BEGIN
    for (int g = 0; g < (int)s->bits.size(); g++) {
        const uint64_t *k = keys + offsets[g];
        if (s->bits[(size_t)g] == 32) {
            std::vector<uint32_t> key((size_t)lengths[g]);
            for (int j = 0; j < lengths[g]; j++)
                key[(size_t)j] = (uint32_t)k[j];
            mtgp32_fast_t tmp;
            if (mtgp32_init_by_array(&tmp, &s->p32[(size_t)g], key.data(), lengths[g]))
                throw Error("out of host memory");
            upload_window(s, g, tmp.status->array);
            mtgp32_free(&tmp);
        } else {
            std::vector<uint64_t> key(k, k + lengths[g]);
            mtgp64_fast_t tmp;
            if (mtgp64_init_by_array(&tmp, &s->p64[(size_t)g], key.data(), lengths[g]))
                throw Error("out of host memory");
            upload_window(s, g, tmp.status->array);
            mtgp64_free(&tmp);
        }
    }
    return 0;
    API_END(-1)
}

int mtgp_genset_set_state(mtgp_genset_t *s, int g, const void *w) {
    API_BEGIN
    if (g < 0 || g >= (int)s->bits.size())
        throw Error("generator index out of range");
    upload_window(s, g, w);
    return 0;
    API_END(-1)
}

int mtgp_genset_get_state(mtgp_genset_t *s, int g, void *w) {
    API_BEGIN
    if (g < 0 || g >= (int)s->bits.size())
        throw Error("generator index out of range");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    hip_check(hipMemcpy(w, s->state(g, s->cur[(size_t)g]), (size_t)s->n[(size_t)g] * (size_t)(s->bits[(size_t)g] / 8),
                        hipMemcpyDeviceToHost),
              "hipMemcpy state");
    return 0;
    API_END(-1)
}

mtgp_plan_t *mtgp_plan_create(mtgp_genset_t *s, const mtgp_request_t *reqs, int nreq, uint64_t segment) {
    API_BEGIN
    std::unique_ptr<mtgp_plan> p(new mtgp_plan());
    p->set = s;
    p->plan.build(*s->ctx, s->refs(), std::vector<mtgp_request_t>(reqs, reqs + nreq), segment);
    p->record_targets();
    return p.release();
    API_END(nullptr)
}

int mtgp_plan_launch(mtgp_plan_t *p, void *stream) {
    API_BEGIN
    p->plan.launch((hipStream_t)stream);
    p->launched = true;
    return 0;
    API_END(-1)
}

int mtgp_plan_commit(mtgp_plan_t *p) {
    API_BEGIN
    const auto &cm = p->plan.commits;
    if (cm.empty())
        return 0;  // nothing moves, or already committed
    if (!p->launched)
        throw Error("plan was never launched since it was built or rebased: its states are not written");
    for (size_t i = 0; i < cm.size(); i++)
        if (p->set->epoch[(size_t)cm[i]] != p->epochs[i])
            throw Error("generator " + std::to_string(cm[i]) +
                        ": state changed since the plan was built or rebased (another commit, a jump, "
                        "set_state, init or substreams); rebase the plan and launch it again");
    for (size_t i = 0; i < cm.size(); i++)
        p->set->set_cur(cm[i], p->out_slot[i]);
    p->plan.commits.clear();
    return 0;
    API_END(-1)
}

int mtgp_plan_rebase(mtgp_plan_t *p, void *stream) {
    API_BEGIN
    p->plan.retarget(p->set->refs(), (hipStream_t)stream);
    p->record_targets();
    return 0;
    API_END(-1)
}

void mtgp_plan_destroy(mtgp_plan_t *p) { delete p; }

int mtgp_plan_launch_part(mtgp_plan_t *p, int part, void *stream) {
    API_BEGIN
    if (part != 0 && part != 1)
        throw Error("part must be 0 (start-state jumps) or 1 (generation)");
    p->plan.launch_part(part, (hipStream_t)stream);
    if (part == 1)
        p->launched = true;
    return 0;
    API_END(-1)
}

int mtgp_plan_configure(mtgp_plan_t *p, int layout, int blocks_per_cu) {
    API_BEGIN
    p->plan.configure(layout, blocks_per_cu);
    return 0;
    API_END(-1)
}

int mtgp_plan_set_overlap(mtgp_plan_t *p, int on) {
    API_BEGIN
    p->plan.set_overlap(on != 0);
    return p->plan.two_phase() ? 1 : 0;
    API_END(-1)
}

int mtgp_plan_capture(mtgp_plan_t *p, void *stream) {
    API_BEGIN
    p->plan.capture((hipStream_t)stream);
    return 0;
    API_END(-1)
}

int mtgp_plan_launch_graph(mtgp_plan_t *p, void *stream) {
    API_BEGIN
    p->plan.launch_graph((hipStream_t)stream);
    p->launched = true;
    return 0;
    API_END(-1)
}

int mtgp_plan_autotune(mtgp_plan_t *p, void *stream, int reps, int *layout, int *blocks_per_cu, float *best_ms) {
    API_BEGIN
    const float ms = p->plan.autotune((hipStream_t)stream, reps, layout, blocks_per_cu);
    p->launched = true;  // its trials wrote the outputs and next states
    if (best_ms)
        *best_ms = ms;
    return 0;
    API_END(-1)
}

int mtgp_plan_stats(const mtgp_plan_t *p, int *items, int *launches, double *plan_ms) {
    if (items)
        *items = (int)p->plan.items.size();
    if (launches) {  // kernel launches of one mtgp_plan_launch
        const Plan &P = p->plan;
        int m = 0;
        for (int ph = 0; ph < 3; ph++) {
            if ((ph == 2) == P.two_phase() || P.mrange[ph].g1 <= P.mrange[ph].g0)
                continue;
            m++;
            for (int rt = 1; rt <= kMjRowTiles; rt++)
                m += P.mrange[ph].i1[rt] > P.mrange[ph].i0[rt];
        }
        *launches = m + (P.two_phase() ? (int)(P.launches.size() + P.jlaunches.size())
                                       : (int)(P.mlaunches.size() + P.mjlaunches.size()));
    }
    if (plan_ms)
        *plan_ms = p->plan.plan_ms;
    return 0;
}

int mtgp_plan_jump_stats(const mtgp_plan_t *p, int *jumps, int *matrix_jumps, int *matrix_groups) {
    if (jumps)
        *jumps = p->plan.jumps;
    if (matrix_jumps)
        *matrix_jumps = p->plan.mjumps;
    if (matrix_groups)
        *matrix_groups = (int)p->plan.mgroups.size();
    return 0;
}

int mtgp_plan_launch_info(const mtgp_plan_t *p, int k, int *bits, int *cls, int *layout, int *grid, int *items,
                          char *kernel, size_t cap) {
    API_BEGIN
    const auto &L = p->plan.mlaunches;
    if (k < 0)
        return (int)L.size();
    if (k >= (int)L.size())
        throw Error("launch index out of range");
    const Plan::Launch &l = L[(size_t)k];
    if (bits)
        *bits = l.bits;
    if (cls)
        *cls = l.cls;
    if (layout)
        *layout = l.layout;
    if (grid)
        *grid = l.grid;
    if (items)
        *items = l.count;
    if (kernel && put_string(kernel_name(l.bits, l.cls, l.layout), kernel, cap) < 0)
        return -1;
    return (int)L.size();
    API_END(-1)
}

int mtgp_genset_generate(mtgp_genset_t *s, const mtgp_request_t *reqs, int nreq, void *stream) {
    API_BEGIN
    std::lock_guard<std::mutex> lk(s->gen_mu);
    if (nreq < 0 || (nreq > 0 && !reqs))
        throw Error("bad request list");
    if (!s->gen_plan)
        s->gen_plan = std::make_unique<Plan>();
    Plan &plan = *s->gen_plan;
    std::vector<mtgp_request_t> rv(reqs, reqs + nreq);
    uint64_t epoch;
    {
        std::shared_lock<std::shared_mutex> tune_lock(s->ctx->tune_mu);
        epoch = s->ctx->tune_epoch;
    }
    auto same = [&]() {
        if (rv.empty() || epoch != s->gen_epoch || rv.size() != s->gen_reqs.size())
            return false;
        for (size_t i = 0; i < rv.size(); i++) {
            const mtgp_request_t &a = rv[i], &b = s->gen_reqs[i];
            if (a.gen != b.gen || a.kind != b.kind || a.count != b.count || a.skip != b.skip || a.out != b.out)
                return false;
        }
        return true;
    };
    if (same()) {
        plan.retarget(s->refs(), (hipStream_t)stream);
    } else {
        s->gen_reqs.clear();  // a failed build leaves nothing to reuse
        s->gen_epoch = ~0ull;
        plan.build(*s->ctx, s->refs(), rv, 0);
        s->gen_reqs = rv;
        s->gen_epoch = epoch;
    }
    plan.launch((hipStream_t)stream);
    hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
    for (int g : plan.commits)
        s->set_cur(g, s->cur[(size_t)g] ^ 1);
    return 0;
    API_END(-1)
}

int mtgp_genset_jump(mtgp_genset_t *s, int g, const char *hex) {
    API_BEGIN
    if (g < 0 || g >= (int)s->bits.size())
        throw Error("generator index out of range");
    const int cur = s->cur[(size_t)g];
    device_jump(*s->ctx, s->slot[(size_t)g], s->bits[(size_t)g], s->mexp[(size_t)g], s->state(g, cur),
                s->state(g, cur ^ 1), hex, false);
    s->set_cur(g, cur ^ 1);
    return 0;
    API_END(-1)
}

int mtgp_genset_substreams(mtgp_genset_t *s, const void *base, const char *hex) {
    API_BEGIN
    const int G = (int)s->bits.size();
    for (int g = 1; g < G; g++)
        if (s->slot[(size_t)g] != s->slot[0])
            throw Error("substreams need one parameter set for all generators");
    if (!hex || !base)
        throw Error("NULL base window or jump string");
    Context &c = *s->ctx;
    const int bits = s->bits[0], cls = c.class_of(s->slot[0]);
    const size_t wbytes = (size_t)s->n[0] * (size_t)(bits / 8);
    const size_t stride = align256(wbytes);
    const int occ = c.jump_blocks_per_cu[bits == 64][cls];
    auto &W = s->sub;
    // Many substreams of an MTGP32 set: every one is a jump from the base,
    // so all of them are one matrix-core group (J^t for t < G from the
    // host, once per jump string).
    const bool matrix = bits == 32 && c.jump_mfma_min > 0 && G - 1 >= c.jump_mfma_min;
    if (matrix && (!W.d || W.key != hex || !W.matrix)) {
        gf2::Poly J;
        if (!gf2::from_hex(hex, J))
            throw Error("jump string contains a non-hex character");
        auto mod = modulus_for(c, s->slot[0]);
        J = mod->reduce_any(J);
        const int nj = G - 1, njp = (nj + 31) / 32 * 32;
        const int kpad = mj_kpad(s->mexp[0]), wc = kpad / 32, pw = mj_pw(kpad, s->n[0]);
        const int pwords = (s->mexp[0] + 31) / 32;
        // coef[w * njp + (t - 1)] = word w of J^t, chains of J products in
        // blocks on the host pool (a block starts from J^t0 by squarings)
        std::vector<uint32_t> coef((size_t)wc * (size_t)njp, 0u);
        constexpr int kBlock = 128;
        parallel_for((nj + kBlock - 1) / kBlock, [&](int blk) {
            const int t0 = 1 + blk * kBlock, t1 = std::min(G, t0 + kBlock);
            gf2::Poly cur, sq = J;
            for (int e = t0; e; e >>= 1) {
                if (e & 1)
                    cur = cur.empty() ? sq : mod->mulmod(cur, sq);
                if (e >> 1)
                    sq = mod->sqrmod(sq);
            }
            for (int t = t0; t < t1; t++) {
                if (t > t0)
                    cur = mod->mulmod(cur, J);
                const std::vector<uint32_t> w = gf2::to_words32(cur, pwords);
                for (size_t x = 0; x < w.size() && x < (size_t)wc; x++)
                    coef[x * (size_t)njp + (size_t)(t - 1)] = w[x];
            }
        });
        // workspace: [base | group | items | coef | planes | windows 1..G-1 | combine items]
        std::vector<MJumpItem> mit[kMjRowTiles + 1];
        for (int r0 = 0; r0 < nj; r0 += 32 * kMjRowTiles) {
            const int rt = std::min(kMjRowTiles, (nj - r0 + 31) / 32);
            for (int j0 = 0; j0 < s->n[0]; j0 += mj_words(rt))
                mit[rt].push_back(MJumpItem{0, r0, j0, 0});
        }
        std::vector<MJumpItem> items;
        for (int rt = 1; rt <= kMjRowTiles; rt++) {
            W.mi_first[rt] = (int)items.size();
            W.mi_count[rt] = (int)mit[rt].size();
            items.insert(items.end(), mit[rt].begin(), mit[rt].end());
        }
        W.o_mgroup = stride;
        W.o_mitems = align256(W.o_mgroup + sizeof(MJumpGroup));
        const size_t o_coef = align256(W.o_mitems + items.size() * sizeof(MJumpItem));
        const size_t o_planes = align256(o_coef + coef.size() * 4);
        const size_t o_out = align256(o_planes + (size_t)pw * 32 * 4);
        W.o_items = align256(o_out + (size_t)nj * wbytes);
        const size_t need = W.o_items + (size_t)G * sizeof(JumpItem);
        W.key.clear();
        if (need > W.cap) {
            if (W.d)
                hip_check(hipFree(W.d), "hipFree");
            W.d = nullptr;
            W.cap = 0;
            hip_check(hipMalloc((void **)&W.d, need), "hipMalloc substream workspace");
            W.cap = need;
        }
        const uint64_t dw = (uint64_t)(uintptr_t)W.d;
        MJumpGroup gr{};
        gr.state_in = dw;
        gr.in_parts = 1;
        gr.planes = dw + o_planes;
        gr.coef = dw + o_coef;
        gr.out = dw + o_out;
        gr.out_stride = wbytes;
        gr.nj = nj;
        gr.njp = njp;
        gr.kpad = kpad;
        gr.pw = pw;
        gr.param = s->slot[0];
        gr.n = s->n[0];
        // combine: substream t's window (the base for t = 0) into slot 0
        std::vector<JumpItem> copies((size_t)G);
        for (int t = 0; t < G; t++) {
            JumpItem &ji = copies[(size_t)t];
            ji.state_in = t ? dw + o_out + (uint64_t)(t - 1) * wbytes : dw;
            ji.in_parts = 1;
            ji.states_out = (uint64_t)(uintptr_t)s->state(t, 0);
            ji.state_bytes = (int32_t)wbytes;
        }
        W.combine_first = 0;
        W.combine_count = G;
        hip_check(hipMemcpy(W.d + W.o_mgroup, &gr, sizeof gr, hipMemcpyHostToDevice), "hipMemcpy group");
        hip_check(hipMemcpy(W.d + W.o_mitems, items.data(), items.size() * sizeof(MJumpItem), hipMemcpyHostToDevice),
                  "hipMemcpy items");
        hip_check(hipMemcpy(W.d + o_coef, coef.data(), coef.size() * 4, hipMemcpyHostToDevice), "hipMemcpy coef");
        hip_check(hipMemcpy(W.d + W.o_items, copies.data(), copies.size() * sizeof(JumpItem), hipMemcpyHostToDevice),
                  "hipMemcpy combine items");
        W.first.clear();
        W.matrix = true;
        W.key = hex;
    }
    if (matrix) {
        hip_check(hipMemcpy(W.d, base, wbytes, hipMemcpyHostToDevice), "hipMemcpy base");
        const MJumpGroup *d_group = reinterpret_cast<const MJumpGroup *>(W.d + W.o_mgroup);
        const MJumpItem *d_mitems = reinterpret_cast<const MJumpItem *>(W.d + W.o_mitems);
        hip_check(launch_planes32(d_group, 1, c.d_params, 0), "planes kernel launch");
        for (int rt = 1; rt <= kMjRowTiles; rt++)
            hip_check(launch_jump_mfma32(d_group, d_mitems + W.mi_first[rt], W.mi_count[rt], rt, 0),
                      "matrix-core jump launch");
        hip_check(launch_combine(reinterpret_cast<const JumpItem *>(W.d + W.o_items), G, 0), "combine launch");
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        for (int g = 0; g < G; g++)
            s->set_cur(g, 0);
        return 0;
    }
    if (!W.d || W.key != hex || W.matrix) {
        W.matrix = false;
        // Substream t = J^t applied to the base, built on the device in
        // radix-4 levels: level m makes t = a 4^m + j (a = 1..3, j < 4^m)
        // as J^(a 4^m) applied to substream j -- the decomposition of t in
        // powers of J the reference's substream samples use over their
        // 4^k 3^162 tables (openCL-sample/mtgp32-jump.cl:697-731).  The host
        // forms the three polynomials of each level by squarings and one
        // product; each level is one batched jump launch.  Small levels
        // split every jump into partial jumps (chunk ranges) so the launch
        // still spreads over the chip; a level reads such a window as the
        // XOR of its parts (JumpItem.in_parts), and a last launch XORs every
        // substream's parts into the set's slot 0.
        gf2::Poly J;
        if (!gf2::from_hex(hex, J))
            throw Error("jump string contains a non-hex character");
        auto mod = modulus_for(c, s->slot[0]);
        J = mod->reduce_any(J);
        const int pw = (s->mexp[0] + 31) / 32;
        constexpr int R = 4;
        std::vector<long> pow_r{1};  // R^m
        while (pow_r.back() < G)
            pow_r.push_back(pow_r.back() * R);
        const int levels = (int)pow_r.size() - 1;
        std::vector<uint64_t> masks;
        std::vector<std::array<int, R>> mask_off((size_t)levels), chunks((size_t)levels);
        {
            gf2::Poly Jm = J;  // J^(R^m)
            for (int m = 0; m < levels; m++) {
                gf2::Poly Pa = Jm;
                for (int a = 1; a < R; a++) {
                    if (a > 1)
                        Pa = mod->mulmod(Pa, Jm);
                    const std::vector<uint32_t> w = gf2::to_words32(Pa, pw);
                    mask_off[(size_t)m][(size_t)a] = (int)masks.size();
                    chunks[(size_t)m][(size_t)a] = append_jump_masks({&w}, masks);
                }
                Jm = mod->mulmod(Pa, Jm);  // J^(R^(m+1))
            }
        }
        // Arena: substream t as parts[t] partial windows.
        std::vector<int> parts((size_t)G, 1);
        std::vector<size_t> arena_off((size_t)G, 0);
        size_t arena = 0;
        for (int m = 0; m < levels; m++) {
            const long lo = pow_r[(size_t)m], hi = std::min<long>(pow_r[(size_t)m + 1], G);
            const int n_items = (int)(hi - lo);
            const int want = (c.cus * c.jump_waves_per_cu + n_items - 1) / n_items;
            const int fit = std::max(1, c.cus * occ / n_items);
            const int P = std::max(1, std::min({want, fit, chunks[(size_t)m][1], c.jump_max_parts}));
            for (long t = lo; t < hi; t++) {
                parts[(size_t)t] = P;
                arena_off[(size_t)t] = arena;
                arena += (size_t)P * stride;
            }
        }
        // Addresses are offsets into the workspace until it is allocated.
        constexpr uint64_t kArena = 1ull << 62, kMasks = 1ull << 61;
        auto in_of = [&](long t, JumpItem &ji) {
            if (t == 0) {
                ji.state_in = 0;  // the base window at the workspace start
                ji.in_parts = 1;
            } else {
                ji.state_in = kArena | arena_off[(size_t)t];
                ji.in_parts = parts[(size_t)t];
                ji.in_part_bytes = stride;
            }
        };
        std::vector<JumpItem> items;
        W.first.clear();
        for (int m = 0; m < levels; m++) {
            W.first.push_back((int)items.size());
            const long lo = pow_r[(size_t)m], hi = std::min<long>(pow_r[(size_t)m + 1], G);
            for (long t = lo; t < hi; t++) {
                const int a = (int)(t / lo);
                const long j = t - a * lo;
                const int ch = chunks[(size_t)m][(size_t)a];
                for (int q = 0; q < parts[(size_t)t]; q++) {
                    JumpItem ji{};
                    in_of(j, ji);
                    ji.masks = kMasks | ((uint64_t)mask_off[(size_t)m][(size_t)a] * 8);
                    ji.states_out = kArena | (arena_off[(size_t)t] + (size_t)q * stride);
                    ji.nb = 1;
                    ji.c0 = (int32_t)((long)ch * q / parts[(size_t)t]);
                    ji.c1 = (int32_t)((long)ch * (q + 1) / parts[(size_t)t]);
                    ji.param = s->slot[0];
                    ji.state_bytes = (int32_t)wbytes;
                    items.push_back(ji);
                }
            }
        }
        W.first.push_back((int)items.size());
        W.combine_first = (int)items.size();
        for (long t = 0; t < G; t++) {
            JumpItem ji{};
            in_of(t, ji);
            ji.states_out = (uint64_t)(uintptr_t)s->state((int)t, 0);
            ji.state_bytes = (int32_t)wbytes;
            items.push_back(ji);
        }
        W.combine_count = G;
        const size_t o_masks = stride;
        W.o_items = align256(o_masks + masks.size() * 8);
        W.o_queue = align256(W.o_items + items.size() * sizeof(JumpItem));
        const size_t o_arena = align256(W.o_queue + 2 * sizeof(uint32_t) * (size_t)levels);
        const size_t need = o_arena + arena;
        W.key.clear();
        if (need > W.cap) {
            if (W.d)
                hip_check(hipFree(W.d), "hipFree");
            W.d = nullptr;
            W.cap = 0;
            hip_check(hipMalloc((void **)&W.d, need), "hipMalloc substream workspace");
            W.cap = need;
        }
        const uint64_t dw = (uint64_t)(uintptr_t)W.d;
        for (auto &ji : items) {
            ji.state_in = (ji.state_in & kArena) ? dw + o_arena + (ji.state_in & ~kArena) : dw;
            if (ji.masks & kMasks)
                ji.masks = dw + o_masks + (ji.masks & ~kMasks);
            if (ji.states_out & kArena)
                ji.states_out = dw + o_arena + (ji.states_out & ~kArena);
        }
        if (!masks.empty())
            hip_check(hipMemcpy(W.d + o_masks, masks.data(), masks.size() * 8, hipMemcpyHostToDevice),
                      "hipMemcpy masks");
        hip_check(hipMemcpy(W.d + W.o_items, items.data(), items.size() * sizeof(JumpItem), hipMemcpyHostToDevice),
                  "hipMemcpy items");
        hip_check(hipMemset(W.d + W.o_queue, 0, 2 * sizeof(uint32_t) * (size_t)levels), "hipMemset queue");
        W.key = hex;
    }
    hip_check(hipMemcpy(W.d, base, wbytes, hipMemcpyHostToDevice), "hipMemcpy base");
    const JumpItem *d_items = reinterpret_cast<const JumpItem *>(W.d + W.o_items);
    for (size_t k = 0; k + 1 < W.first.size(); k++) {
        const int f = W.first[k], n_items = W.first[k + 1] - f;
        if (!n_items)
            continue;
        hip_check(launch_jumps(bits, cls, d_items + f, n_items, c.d_params,
                               reinterpret_cast<uint32_t *>(W.d + W.o_queue) + 2 * k,
                               std::min(n_items, c.cus * occ), 0, 0),
                  "jump kernel launch");
    }
    hip_check(launch_combine(d_items + W.combine_first, W.combine_count, 0), "combine launch");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    for (int g = 0; g < G; g++)
        s->set_cur(g, 0);
    return 0;
    API_END(-1)
}

// ------------------------------------------- host-struct bulk generation

int mtgp32_fill(mtgp32_fast_t *g, int kind, void *out, size_t n) {
    API_BEGIN
    return fill_impl<mtgp32_fast_t, uint32_t>(g, kind, out, n, 32);
    API_END(-1)
}

int mtgp64_fill(mtgp64_fast_t *g, int kind, void *out, size_t n) {
    API_BEGIN
    if (g)
        check64(g->params);
    return fill_impl<mtgp64_fast_t, uint64_t>(g, kind, out, n, 64);
    API_END(-1)
}

// ---------------------------------------------- jump drop-ins and polys

// Errors cannot be returned through the reference signature: the state is
// then left unchanged and mtgp_last_error() is non-empty ("" on success).
void mtgp32_fast_jump(mtgp32_fast_t *g, const char *hex) {
    try {
        fast_jump_impl<mtgp32_fast_t, uint32_t>(g, hex, 32);
        set_error("");
    } catch (const std::exception &e) {
        set_error(e.what());
    }
}

void mtgp64_fast_jump(mtgp64_fast_t *g, const char *hex) {
    try {
        check64(g->params);
        fast_jump_impl<mtgp64_fast_t, uint64_t>(g, hex, 64);
        set_error("");
    } catch (const std::exception &e) {
        set_error(e.what());
    }
}

int mtgp32_charpoly_hex(const mtgp32_params_fast_t *p, char *buf, size_t cap) {
    API_BEGIN
    return charpoly_impl(*p, buf, cap);
    API_END(-1)
}

int mtgp64_charpoly_hex(const mtgp64_params_fast_t *p, char *buf, size_t cap) {
    API_BEGIN
    check64(*p);
    return charpoly_impl(*p, buf, cap);
    API_END(-1)
}

int mtgp32_jump_hex(const mtgp32_params_fast_t *p, const uint64_t *steps, int nwords, char *buf, size_t cap) {
    API_BEGIN
    return jump_impl(*p, steps, nwords, buf, cap);
    API_END(-1)
}

int mtgp64_jump_hex(const mtgp64_params_fast_t *p, const uint64_t *steps, int nwords, char *buf, size_t cap) {
    API_BEGIN
    check64(*p);
    return jump_impl(*p, steps, nwords, buf, cap);
    API_END(-1)
}

}  // extern "C"
