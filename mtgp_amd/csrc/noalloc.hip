// noalloc.hip -- mtgp32_fast_jump_noalloc: the GPU jump of mtgp32_fast_jump
// with every piece of device memory it needs taken from one static device
// array, for callers that must not allocate (code under graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// mtgp32_fast_jump (runtime.cpp) allocates, per call, device memory for the
// window in and out, the per-residue coefficient masks, the partial windows
// of the split jump, the JumpItems and the work-queue counters of the
// batched jump kernel.  Here those live in
// `scratch`, a file-scope __device__ array of kScratchBytes: each GPU holds
// its own copy from module load until the process ends.  Host code never
// passes the array's name to the runtime (on the host it names a stand-in,
// not device memory): the first call on a device looks its device address
// up with hipGetSymbolAddress and every later call on that device uses that
// address.  Calls that overlap on one GPU share the array and race.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <mutex>

#include "kernels.hpp"
#include "runtime.hpp"

using namespace mtgp;

namespace {

// Largest MTGP32 window (MEXP 44497: N = 1391 words).
constexpr int kMaxWindow = 1391;
// Longest jump string: kMaxChunks chunks of kJumpChunk coefficients
// (16384 hex characters, enough for x^s mod P of every MTGP32 table).
constexpr int kMaxChunks = 64;
// A jump is split into at most kMaxParts partial jumps over chunk ranges
// (one wave each), whose windows a second launch XORs together.
constexpr int kMaxParts = 16;
constexpr size_t kWinBytes = ((size_t)kMaxWindow * 4 + 255) & ~(size_t)255;
constexpr size_t kQueueOff = 0;                                          // 2 pairs of uint32_t
constexpr size_t kItemOff = 256;                                         // kMaxParts + 1 JumpItems
constexpr size_t kParamOff = kItemOff + (((kMaxParts + 1) * 64 + 255) & ~(size_t)255);  // DevParam
constexpr size_t kIdMaskOff = kParamOff + 256;                           // identity masks (1 chunk)
constexpr size_t kMaskOff = kIdMaskOff + 64 * 8;                         // chunks x 64 x uint64_t
constexpr size_t kWinInOff = kMaskOff + (size_t)kMaxChunks * 64 * 8;     // window in
constexpr size_t kPartsOff = kWinInOff + kWinBytes;                      // partial windows
constexpr size_t kWinOutOff = kPartsOff + (size_t)kMaxParts * kWinBytes;  // window out
constexpr size_t kScratchBytes = kWinOutOff + kWinBytes;
static_assert(sizeof(JumpItem) == 64 && sizeof(DevParam) <= kIdMaskOff - kParamOff, "scratch layout");

}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t *g_scratch[kMaxDevices];  // device address of `scratch` per device

// Device address of this device's copy of `scratch`; looked up once per
// device with that device current.
uint8_t *scratch_ptr(int dev) {
    if (dev < 0 || dev >= kMaxDevices)
        throw Error("device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr)
            throw Error("hipGetSymbolAddress(scratch) returned NULL");
        g_scratch[dev] = static_cast<uint8_t *>(ptr);
    }
    return g_scratch[dev];
}

int hex_value(char ch) {
    if (ch >= '0' && ch <= '9')
        return ch - '0';
    if (ch >= 'a' && ch <= 'f')
        return ch - 'a' + 10;
    if (ch >= 'A' && ch <= 'F')
        return ch - 'A' + 10;
    return -1;
}

void jump_noalloc(mtgp32_fast_t *g, const char *hex) {
    if (!g || !g->status)
        throw Error("generator is not initialised");
    if (!hex)
        throw Error("NULL jump string");
    const int N = g->status->size;
    const DevParam rec = dev_param_of(g->params);
    if (N != rec.n || N > kMaxWindow)
        throw Error("state size does not match the parameter set");
    const int cls = kernel_class(32, rec.n, rec.pos);
    if (cls < 0)
        throw Error("parameter set fits no device jump kernel");
    const size_t len = strlen(hex);
    // mtgp32_fast_jump applies every coefficient of the string (at least one
    // 32-coefficient word), so the kernel folds that many chunks.
    const size_t ncoeff = len * 4 > 32 ? len * 4 : 32;
    const int chunks = (int)((ncoeff + kJumpChunk - 1) / kJumpChunk);
    if (chunks > kMaxChunks)
        throw Error("jump string longer than the no-allocation limit of " +
                    std::to_string(kMaxChunks * kJumpChunk / 4) + " hex characters");
    // Masks (layout of JumpItem.masks, one jump in bits 0..15): bit q of
    // word c * 64 + rho is the coefficient of x^(c * 1024 + 64 q + rho).
    uint64_t masks[(size_t)kMaxChunks * 64];
    memset(masks, 0, (size_t)chunks * 64 * sizeof(uint64_t));
    for (size_t i = 0; i < len; i++) {
        const int v = hex_value(hex[i]);
        if (v < 0)
            throw Error("jump string contains a non-hex character");
        for (int t = 0; t < 4; t++)
            if (v >> t & 1) {
                const size_t k = 4 * i + (size_t)t;
                masks[(k / kJumpChunk) * 64 + k % 64] |= 1ull << ((k % kJumpChunk) / 64);
            }
    }
    uint32_t w[kMaxWindow];
    const mtgp32_status_fast_t *st = g->status;
    for (int j = 0; j < N; j++)
        w[j] = st->array[(st->idx - N + 1 + j + st->large_size) & st->large_mask];

    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    uint8_t *base = scratch_ptr(dev);
    const int parts = chunks < kMaxParts ? chunks : kMaxParts;
    JumpItem ji[kMaxParts + 1];
    memset(ji, 0, sizeof ji);
    for (int q = 0; q < parts; q++) {
        ji[q].state_in = (uint64_t)(uintptr_t)(base + kWinInOff);
        ji[q].masks = (uint64_t)(uintptr_t)(base + kMaskOff);
        ji[q].states_out = (uint64_t)(uintptr_t)(base + kPartsOff + (size_t)q * kWinBytes);
        ji[q].nb = 1;
        ji[q].c0 = chunks * q / parts;
        ji[q].c1 = chunks * (q + 1) / parts;
        ji[q].param = 0;
        ji[q].state_bytes = N * 4;
    }
    // the identity polynomial (coefficient 0) over the XOR of the parts
    JumpItem &cj = ji[parts];
    cj.state_in = (uint64_t)(uintptr_t)(base + kPartsOff);
    cj.in_parts = parts;
    cj.in_part_bytes = kWinBytes;
    cj.masks = (uint64_t)(uintptr_t)(base + kIdMaskOff);
    cj.states_out = (uint64_t)(uintptr_t)(base + kWinOutOff);
    cj.nb = 1;
    cj.c0 = 0;
    cj.c1 = 1;
    cj.param = 0;
    cj.state_bytes = N * 4;
    uint64_t id_masks[64];
    memset(id_masks, 0, sizeof id_masks);
    id_masks[0] = 1;
    hip_check(hipMemcpy(base + kWinInOff, w, (size_t)N * 4, hipMemcpyHostToDevice), "hipMemcpy window");
    hip_check(hipMemcpy(base + kMaskOff, masks, (size_t)chunks * 64 * sizeof(uint64_t), hipMemcpyHostToDevice),
              "hipMemcpy masks");
    hip_check(hipMemcpy(base + kIdMaskOff, id_masks, sizeof id_masks, hipMemcpyHostToDevice), "hipMemcpy masks");
    hip_check(hipMemcpy(base + kItemOff, ji, (size_t)(parts + 1) * sizeof(JumpItem), hipMemcpyHostToDevice),
              "hipMemcpy items");
    hip_check(hipMemcpy(base + kParamOff, &rec, sizeof(rec), hipMemcpyHostToDevice), "hipMemcpy param");
    // The kernels' work-queue pairs must be zero on entry (a launch leaves
    // its pair zero, but one that failed part-way may not have).
    hip_check(hipMemsetAsync(base + kQueueOff, 0, 4 * sizeof(uint32_t), 0), "hipMemsetAsync queue");
    const JumpItem *items = reinterpret_cast<const JumpItem *>(base + kItemOff);
    const DevParam *param = reinterpret_cast<const DevParam *>(base + kParamOff);
    uint32_t *queue = reinterpret_cast<uint32_t *>(base + kQueueOff);
    hip_check(launch_jumps(32, cls, items, parts, param, queue, parts, 0, 0), "jump kernel launch");
    hip_check(launch_jumps(32, cls, items + parts, 1, param, queue + 2, 1, 0, 0), "jump kernel launch");
    hip_check(hipMemcpy(w, base + kWinOutOff, (size_t)N * 4, hipMemcpyDeviceToHost), "hipMemcpy window back");
    memcpy(g->status->array, w, (size_t)N * 4);
    g->status->idx = N - 1;
}

}  // namespace

extern "C" int mtgp32_fast_jump_noalloc(mtgp32_fast_t *mtgp32, const char *jump_string) {
    try {
        jump_noalloc(mtgp32, jump_string);
        set_error("");
        return 0;
    } catch (const std::exception &e) {
        set_error(e.what());
        return -1;
    }
}
