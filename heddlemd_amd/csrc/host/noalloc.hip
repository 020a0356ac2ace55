// Pair snapshot into a fixed device scratch array: the variant of
// hmd_nbl_pair_snapshot for callers that cannot allocate device memory.
//
// hmd_nbl_pair_snapshot stages the listed pairs and their counter in device
// memory it allocates per call.  Here both live in `scratch`, a file-scope
// device array of kScratchBytes per GPU (held until the process ends): a
// 256-byte header whose first 8 bytes are the pair counter, then room for
// kMaxPairs (i, j) pairs.  Its device address is looked up once per device
// (hipGetSymbolAddress, with that device current) and only that address is
// handed to kernels and copies.  Every call on one GPU shares the array and
// starts by clearing the counter; calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "forcefield.h"

namespace {

constexpr uint64_t kMaxPairs = 1ull << 21;               // 2,097,152 pairs
constexpr size_t kHeaderBytes = 256;
constexpr size_t kScratchBytes = kHeaderBytes + kMaxPairs * 2 * sizeof(uint32_t);

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on `device` (made current by the caller).
unsigned char* scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices)
    throw hmd::Error(HMD_ERR_INVALID, "device ordinal out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[device]) {
    void* ptr = nullptr;
    hmd::check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch[device] = static_cast<unsigned char*>(ptr);
  }
  return g_scratch[device];
}

// Makes `device` current for the scope and puts the caller's back after.
struct DeviceScope {
  int prev = -1;
  explicit DeviceScope(int device) {
    hmd::check(hipGetDevice(&prev), "hipGetDevice");
    hmd::check(hipSetDevice(device), "hipSetDevice");
  }
  ~DeviceScope() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

}  // namespace

using namespace hmd;

extern "C" hmd_status hmd_nbl_pair_snapshot_noalloc(hmd_forcefield* f, uint32_t* pairs,
                                                    uint64_t cap, uint64_t* n_pairs) {
  HMD_GUARD({
    if (!f || !n_pairs || (cap && !pairs)) throw Error(HMD_ERR_INVALID, "null argument");
    if (cap > kMaxPairs)
      throw Error(HMD_ERR_INVALID,
                  "cap " + std::to_string(cap) + " exceeds the no-allocation scratch of " +
                      std::to_string(kMaxPairs) + " pairs");
    ForceField& F = *f->ff;
    Context& c = F.c;
    DeviceScope on(c.device);
    unsigned char* base = scratch_on(c.device);
    auto* count = reinterpret_cast<unsigned long long*>(base);
    auto* staged = reinterpret_cast<uint32_t*>(base + kHeaderBytes);
    check(hipMemsetAsync(count, 0, sizeof(unsigned long long), c.stream), "hipMemsetAsync");
    uint32_t nw = 0;
    F.n_work.download(&nw, 1, c.stream);
    NblPairsParams pp{F.posq_sorted.ptr, F.sorted_id.ptr, F.work.ptr, F.n_work.ptr,
                      F.entry_j.ptr, F.p.lattice, F.n, F.cap_entries, F.r_search_sq,
                      staged, count, cap, nullptr};
    launch(c, K_hmd_nbl_pairs, ceil_div(nw, 4), 256, 0, pp);
    unsigned long long total = 0;
    check(hipMemcpyAsync(&total, count, sizeof total, hipMemcpyDeviceToHost, c.stream), "D2H");
    check(hipStreamSynchronize(c.stream), "sync");
    *n_pairs = total;
    const uint64_t k = std::min<uint64_t>(total, cap);
    if (k) {
      check(hipMemcpyAsync(pairs, staged, 2 * k * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           c.stream),
            "D2H");
      check(hipStreamSynchronize(c.stream), "sync");
    }
  });
}
