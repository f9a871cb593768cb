// Internal helpers shared by the library's HIP and C++ sources.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>

#include "../../include/einfold_amd.h"

namespace ef {

void set_error(const std::string &message);
void clear_error();

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// A HIP call's failure as a status: device memory exhaustion is
// EF_OUT_OF_MEMORY (the caller falls back to the host plan), anything else
// EF_DEVICE_ERROR.
ef_status hip_status(hipError_t e, const char *what);

#define EF_HIP(call)                                             \
    do {                                                         \
        hipError_t ef_e_ = (call);                               \
        if (ef_e_ != hipSuccess) return ef::hip_status(ef_e_, #call); \
    } while (0)

#define EF_TRY(call)                                             \
    do {                                                         \
        ef_status ef_s_ = (call);                                \
        if (ef_s_ != EF_OK) return ef_s_;                        \
    } while (0)

// Device memory owned by a scope.
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
    ~DeviceBuffer() { if (ptr) (void)hipFree(ptr); }
    ef_status alloc(size_t n) {
        if (ptr) { (void)hipFree(ptr); ptr = nullptr; }
        bytes = n;
        if (n == 0) return EF_OK;
        hipError_t e = hipMalloc(&ptr, n);
        if (e != hipSuccess) {
            // Reset the thread's last error, so that a later launch's
            // hipGetLastError() does not report this refusal again.
            (void)hipGetLastError();
            ptr = nullptr;
            return hip_status(e, "hipMalloc");
        }
        return EF_OK;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

inline unsigned blocks_for(size_t n, unsigned threads) {
    size_t b = (n + threads - 1) / threads;
    return (unsigned)(b == 0 ? 1 : (b > 0x7fffffffULL ? 0x7fffffffULL : b));
}

// ---- launchers implemented in the .hip files (host side) ----
ef_status launch_dense_accumulate_f64(const double *a, const uint32_t *held, const double *b,
                                      size_t nx, size_t nz, size_t ny, size_t per_batch,
                                      int predicated, double *out, hipStream_t s);
ef_status launch_dense_product_f32(const float *a, const float *b, float *out, size_t batches,
                                   size_t m, size_t n, size_t k, uint32_t precision,
                                   hipStream_t s);
ef_status launch_dense_product_f64(const double *a, const double *b, double *out, size_t batches,
                                   size_t m, size_t n, size_t k, hipStream_t s);
ef_status rank_u64(const uint64_t *words, size_t n, uint32_t *rank, uint32_t *first,
                   size_t *n_ranks, void *scratch, hipStream_t s);
size_t rank_scratch_bytes(size_t n);
// Stable sort of (word, row) pairs by word; rows_in null means 0..n-1.
// `scratch` is rank_scratch_bytes(n) bytes; w_out may be null.
ef_status sort_u64(const uint64_t *words, const uint32_t *rows_in, size_t n, uint64_t *w_out, uint32_t *r_out,
                   void *scratch, hipStream_t s);
size_t attention_scratch_bytes(const ef_attention_desc &d);
size_t attention_fwd_scratch_bytes(const ef_attention_desc &d);   // the forward pass's share
// This device's address of the fixed scratch array of the no-allocation
// entry point (resolved once per device).
ef_status noalloc_scratch(void **ptr);
// Row statistics are float per query, double for EF_PRECISION_FP64.
ef_status attention_fwd(const ef_attention_desc &d, const void *q, const void *k, const void *v,
                        void *out, void *row_max, void *row_sum, uint64_t *tiles_visited,
                        void *scratch, hipStream_t s);
ef_status attention_bwd(const ef_attention_desc &d, const void *q, const void *k, const void *v,
                        const void *out, const void *dout, const void *row_max,
                        const void *row_sum, void *dq, void *dk, void *dv, void *scratch,
                        int stages, hipStream_t s);   // stages: 1 prepare, 2 dK/dV, 4 dQ

struct GroupedPlan;
ef_status grouped_plan_create(const ef_gemm_problem *problems, size_t count, uint32_t precision, int tile,
                              void *stream, GroupedPlan **out);
ef_status grouped_plan_run(GroupedPlan *plan, void *stream);
int grouped_plan_tile(const GroupedPlan *plan);
void grouped_plan_destroy(GroupedPlan *plan);
// One set of products run once on `s` (no graph), with the autotuner's tile
// shape for its (M, N, K, density) signature; returns after they finish.
// density: per problem, or null for 1.
ef_status grouped_run_once(const ef_gemm_problem *problems, size_t count, uint32_t precision, const double *density,
                           hipStream_t s);

}  // namespace ef
