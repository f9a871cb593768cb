// The attention forward pass for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops, harnesses that
// forbid hipMalloc): ef_attention_fwd with its scratch in a fixed device
// array instead of a buffer of the caller's or the library's.
//
// The array lives in this file; each GPU the library runs on holds its own
// copy, at its own address, until the process ends.  Host code never
// passes the array's name to a kernel or a copy: that is a host-side
// stand-in.  It resolves the device address once per device
// (hipGetSymbolAddress, at ef_init or the first call on a device) and
// passes that address on.
//
// Every call on one GPU shares that GPU's array and leaves its contents
// behind: calls that overlap on one GPU (on different streams, or from
// different threads) race.  What must start at zero (the work counter)
// is cleared by the forward's own prep kernel, not by a memset, so a
// captured graph replays correctly.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "common.h"

namespace {

constexpr size_t kScratchBytes = EF_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex resolve_mu;
void *scratch_ptr[kMaxDevices];

}  // namespace

namespace ef {

ef_status noalloc_scratch(void **ptr) {
    int dev = 0;
    EF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device index past the no-allocation entry point's table");
        return EF_INVALID;
    }
    std::lock_guard<std::mutex> lock(resolve_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return hip_status(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        }
        scratch_ptr[dev] = p;
    }
    *ptr = scratch_ptr[dev];
    return EF_OK;
}

}  // namespace ef

using namespace ef;

extern "C" ef_status ef_attention_fwd_noalloc(const ef_attention_desc *desc, const void *q, const void *k,
                                              const void *v, void *out, void *row_max, void *row_sum,
                                              uint64_t *tiles_visited, void *stream) {
    clear_error();
    if (!desc) {
        set_error("no attention description");
        return EF_INVALID;
    }
    const size_t need = attention_fwd_scratch_bytes(*desc);
    if (need > kScratchBytes) {
        set_error("ef_attention_fwd_noalloc: this description needs " + std::to_string(need) +
                  " bytes of scratch, more than the fixed " + std::to_string(kScratchBytes) +
                  "; call ef_attention_fwd with a scratch buffer instead");
        return EF_INVALID;
    }
    void *sc = nullptr;
    EF_TRY(noalloc_scratch(&sc));
    return attention_fwd(*desc, q, k, v, out, row_max, row_sum, tiles_visited, sc, as_stream(stream));
}
