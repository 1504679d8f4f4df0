// noalloc.hip -- the allocation-free dot product (smg_dot_dev_noalloc).
//
// smg_dot_dev finishes its reduction through the library's per-process partial-sum
// buffer and completion counters, allocated on first use.  This variant keeps its
// partial rows, its completion counter and its result in a file-scope device array,
// `scratch`, so a call allocates nothing: for callers that cannot call hipMalloc
// (code under stream capture, tight serving loops).  The array lives in the code
// object and so exists once per device the library runs on, until the process ends.
//
// Host code never passes `scratch` itself to a kernel: its host-side name stands in
// for the symbol and is not a device address.  The device address of each GPU's copy
// comes from hipGetSymbolAddress, taken once per device with that device current (at
// smg_init, or at the first call on a device) and kept in g_scratch[device].
#include <mutex>

#include "smg_kernels.h"

namespace smg {

// layout: partial rows of at most kMaxRedBlocks workgroups x 2 doubles, then the
// completion counter (its own 256-byte slot), then the {Re, Im} result
constexpr size_t kScratchPart = 2 * (size_t)kMaxRedBlocks;
constexpr size_t kScratchCounter = kScratchPart;         // as unsigned
constexpr size_t kScratchResult = kScratchPart + 32;
constexpr size_t kScratchDoubles = kScratchPart + 64;    // 16.5 KiB
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

__global__ void k_clear_counter(unsigned* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0u;
}
}  // namespace

// the device address of the current device's scratch (looked up once per device)
double* noalloc_scratch() {
    int dev = 0;
    SMG_HIP(hipGetDevice(&dev));
    SMG_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        SMG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        SMG_REQUIRE(p != nullptr, "hipGetSymbolAddress returned a null scratch address");
        g_scratch[dev] = static_cast<double*>(p);
    }
    return g_scratch[dev];
}

// the workgroups launch_dot uses for n entries, and so the partial rows it writes
size_t noalloc_dot_rows(size_t n) { return (size_t)red_blocks(n); }

void noalloc_dot(const void* a, const void* b, size_t n, int f64, double* out2, hipStream_t st) {
    double* base = noalloc_scratch();
    SMG_REQUIRE(2 * noalloc_dot_rows(n) <= kScratchPart, "dot: more partial rows than the scratch holds");
    Fin f;
    f.op = FIN_STORE;
    f.out = base + kScratchResult;
    f.part = base;
    f.counter = reinterpret_cast<unsigned*>(base + kScratchCounter);
    // the counter must start at zero; a kernel (not a memset node) clears it
    k_clear_counter<<<1, 64, 0, st>>>(f.counter);
    SMG_LAUNCH_CHECK();
    if (f64) launch_dot<double>((const c64*)a, (const c64*)b, n, f, st);
    else launch_dot<float>((const c32*)a, (const c32*)b, n, f, st);
    SMG_HIP(hipMemcpyAsync(out2, f.out, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    SMG_HIP(hipStreamSynchronize(st));
}

}  // namespace smg
