// noalloc.hip -- the device-pointer boundary apply for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  Every workspace of the apply (staged
// sources, MFMA operands, column scales, segment partials) lives in one
// file-scope device array, `scratch`, of kScratchBytes; each GPU the entry
// point runs on holds its own copy until the process ends.
//
// The host never uses the array's name as an address: on the first call on a
// device, with that device current, hipGetSymbolAddress gives the device
// address of that device's copy, which is kept per device and passed to the
// kernels.  All calls on one GPU share its array and leave their contents in
// it; calls that overlap on one GPU (different streams or threads) race.
#include <atomic>
#include <mutex>
#include <string>

#include "capi_util.hpp"

namespace {

constexpr size_t kScratchBytes = (size_t)256 << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) char scratch[kScratchBytes];

std::mutex g_scratch_mtx;
std::atomic<char*> g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on `device` (the current device).
char* scratch_on(int device)
{
    if (device < 0 || device >= kMaxDevices) throw baemm::ArgError{"device index beyond the scratch table"};
    char* p = g_scratch_ptr[device].load(std::memory_order_acquire);
    if (p) return p;
    std::lock_guard<std::mutex> lk(g_scratch_mtx);
    p = g_scratch_ptr[device].load(std::memory_order_relaxed);
    if (!p) {
        void* addr = nullptr;
        BAEMM_HIP(hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)));
        if (!addr) throw baemm::ArgError{"hipGetSymbolAddress returned no address for the scratch array"};
        p = static_cast<char*>(addr);
        g_scratch_ptr[device].store(p, std::memory_order_release);
    }
    return p;
}

}  // namespace

using namespace baemm;

extern "C" {

int baemm_apply_boundary_dev_noalloc(baemm_handle h, const double* alpha, const void* B, int64_t ldB, int B_type,
                                     const double* beta, void* C, int64_t ldC, int C_type, void* stream)
{
    return guard(h, [&](Handle& hh) {
        if (!hh.prm.loaded) throw StateError{"parameters are not loaded"};
        if (!B || !C) throw ArgError{"B and C must not be NULL"};
        if ((B_type != kC128 && B_type != kC64) || (C_type != kC128 && C_type != kC64))
            throw ArgError{"unknown element type"};
        if (ldB < hh.prm.W || ldC < hh.prm.W) throw ArgError{"leading dimension smaller than wave_count"};
        char* base = scratch_on(hh.device);
        ApplyArgs a{cplx(alpha, 1.0), cplx(beta, 0.0), B, ldB, B_type, C, ldC, C_type};
        apply_boundary_dev(hh, a, stream_of(hh, stream), base, kScratchBytes);
        return BAEMM_OK;
    });
}

int64_t baemm_noalloc_scratch_bytes(void) { return (int64_t)kScratchBytes; }

}  // extern "C"
