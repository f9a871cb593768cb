// The C ABI's device-resident entry points and the DenseBackend mirror
// (include/einfold_amd.h, levels 2 and 3).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"

namespace ef {

namespace {
thread_local std::string last_error;
}

void set_error(const std::string &message) { last_error = message; }
void clear_error() { last_error.clear(); }

ef_status hip_status(hipError_t e, const char *what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? EF_OUT_OF_MEMORY : EF_DEVICE_ERROR;
}

}  // namespace ef

using namespace ef;

namespace {

// Runs `body` with the caller's scratch, or with a buffer of its own that
// lives until the stream has finished with it.
template <class F>
ef_status with_scratch(const ef_attention_desc *desc, void *scratch, void *stream, F body) {
    if (!desc) {
        set_error("no attention description");
        return EF_INVALID;
    }
    if (scratch) return body(scratch);
    DeviceBuffer own;
    EF_TRY(own.alloc(attention_scratch_bytes(*desc)));
    ef_status st = body(own.ptr);
    EF_HIP(hipStreamSynchronize(as_stream(stream)));
    return st;
}

}  // namespace

extern "C" {

const char *ef_last_error(void) { return last_error.c_str(); }

int ef_version(void) { return 100; }

ef_status ef_init(int device) {
    clear_error();
    EF_HIP(hipSetDevice(device));
    EF_HIP(hipFree(nullptr));
    // This device's address of the no-allocation entry point's array.
    void *fixed = nullptr;
    EF_TRY(noalloc_scratch(&fixed));
    return EF_OK;
}

ef_status ef_dense_accumulate_f64(const double *a, const uint32_t *held, const double *b, size_t nx,
                                  size_t nz, size_t ny, size_t per_batch, double *out, void *stream) {
    clear_error();
    return launch_dense_accumulate_f64(a, held, b, nx, nz, ny, per_batch, 1, out, as_stream(stream));
}

ef_status ef_dense_product_f32(const float *a, const float *b, float *out, size_t batches, size_t m,
                               size_t n, size_t k, uint32_t precision, void *stream) {
    clear_error();
    return launch_dense_product_f32(a, b, out, batches, m, n, k, precision, as_stream(stream));
}

size_t ef_rank_scratch_bytes(size_t n) { return rank_scratch_bytes(n); }

ef_status ef_rank_u64(const uint64_t *words, size_t n, uint32_t *rank, uint32_t *first, size_t *n_ranks,
                      void *scratch, void *stream) {
    clear_error();
    return rank_u64(words, n, rank, first, n_ranks, scratch, as_stream(stream));
}

ef_status ef_grouped_plan_create(const ef_gemm_problem *problems, size_t count, uint32_t precision, int tile,
                                 void *stream, ef_grouped_plan **plan) {
    clear_error();
    if (!plan || (count && !problems)) {
        set_error("problems and a plan to write are required");
        return EF_INVALID;
    }
    GroupedPlan *p = nullptr;
    const ef_status st = grouped_plan_create(problems, count, precision, tile, stream, &p);
    *plan = reinterpret_cast<ef_grouped_plan *>(p);
    return st;
}

ef_status ef_grouped_plan_run(ef_grouped_plan *plan, void *stream) {
    clear_error();
    return grouped_plan_run(reinterpret_cast<GroupedPlan *>(plan), stream);
}

int ef_grouped_plan_tile(const ef_grouped_plan *plan) {
    return grouped_plan_tile(reinterpret_cast<const GroupedPlan *>(plan));
}

void ef_grouped_plan_destroy(ef_grouped_plan *plan) { grouped_plan_destroy(reinterpret_cast<GroupedPlan *>(plan)); }

size_t ef_attention_scratch_bytes(const ef_attention_desc *desc) {
    return desc ? attention_scratch_bytes(*desc) : 0;
}

ef_status ef_attention_fwd(const ef_attention_desc *desc, const void *q, const void *k, const void *v,
                           void *out, void *row_max, void *row_sum, uint64_t *tiles_visited,
                           void *scratch, void *stream) {
    clear_error();
    return with_scratch(desc, scratch, stream, [&](void *sc) {
        return attention_fwd(*desc, q, k, v, out, row_max, row_sum, tiles_visited, sc, as_stream(stream));
    });
}

ef_status ef_attention_bwd(const ef_attention_desc *desc, const void *q, const void *k, const void *v,
                           const void *out, const void *dout, const void *row_max, const void *row_sum,
                           void *dq, void *dk, void *dv, void *scratch, void *stream) {
    clear_error();
    return with_scratch(desc, scratch, stream, [&](void *sc) {
        return attention_bwd(*desc, q, k, v, out, dout, row_max, row_sum, dq, dk, dv, sc, 7, as_stream(stream));
    });
}

ef_status ef_attention_bwd_prepare(const ef_attention_desc *desc, const void *q, const void *k, const void *v,
                                   const void *out, const void *dout, void *scratch, void *stream) {
    clear_error();
    if (!desc || !scratch) {
        set_error("a description and scratch are required");
        return EF_INVALID;
    }
    return attention_bwd(*desc, q, k, v, out, dout, nullptr, nullptr, nullptr, nullptr, nullptr, scratch, 1,
                         as_stream(stream));
}

ef_status ef_attention_bwd_kv(const ef_attention_desc *desc, const void *k, const void *v, const void *row_max,
                              const void *row_sum, void *scratch, void *dk, void *dv, void *stream) {
    clear_error();
    if (!desc || !scratch) {
        set_error("a description and the prepared scratch are required");
        return EF_INVALID;
    }
    return attention_bwd(*desc, nullptr, k, v, nullptr, nullptr, row_max, row_sum, nullptr, dk, dv, scratch, 2,
                         as_stream(stream));
}

ef_status ef_attention_bwd_q(const ef_attention_desc *desc, const void *q, const void *dout, const void *row_max,
                             const void *row_sum, void *scratch, void *dq, void *stream) {
    clear_error();
    if (!desc || !scratch) {
        set_error("a description and the prepared scratch are required");
        return EF_INVALID;
    }
    return attention_bwd(*desc, q, nullptr, nullptr, nullptr, dout, row_max, row_sum, dq, nullptr, nullptr, scratch, 4,
                         as_stream(stream));
}

// ---------------------------------------------------------------------
// DenseBackend (crates/einfold-kernels/src/backend.rs:124-143)
// ---------------------------------------------------------------------

// Whether `bytes` fit the device's free memory with room to spare.
static bool fits_device(size_t bytes) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return false;
    return bytes <= free_b / 10 * 9;
}

int ef_dense_may_run(uint32_t kind, const ef_product_shape *product, const ef_attention_shape *attention) {
    if (kind == EF_WORK_PRODUCT && product) {
        const ef_product_shape &s = *product;
        const double bytes = 8.0 * s.batches * ((double)s.rows * s.inner + (double)s.inner * s.columns +
                                                (double)s.rows * s.columns);
        return s.batches <= 65535 && bytes < 1.0e15 && fits_device((size_t)bytes);
    }
    if (kind == EF_WORK_ATTENTION && attention) {
        const ef_attention_shape &s = *attention;
        const double bytes = 8.0 * s.batches *
                             ((double)s.queries * s.features + (double)s.keys * s.features +
                              (double)s.keys * s.values + (double)s.queries * s.values) +
                             16.0 * s.queries + 8.0 * s.batches * s.queries;
        return s.batches <= 65535 && bytes < 1.0e15 && fits_device((size_t)bytes);
    }
    return 0;
}

ef_status ef_dense_run_product(const double *a, const double *b, ef_product_shape shape, double *out,
                               const ef_options *options) {
    clear_error();
    hipStream_t s = as_stream(options ? options->stream : nullptr);
    const size_t na = shape.batches * shape.rows * shape.inner;
    const size_t nb = shape.batches * shape.inner * shape.columns;
    const size_t nc = shape.batches * shape.rows * shape.columns;
    if (nc == 0) return EF_OK;
    DeviceBuffer da, db, dc;
    EF_TRY(da.alloc(na * 8));
    EF_TRY(db.alloc(nb * 8));
    EF_TRY(dc.alloc(nc * 8));
    if (na) EF_HIP(hipMemcpyAsync(da.ptr, a, na * 8, hipMemcpyHostToDevice, s));
    if (nb) EF_HIP(hipMemcpyAsync(db.ptr, b, nb * 8, hipMemcpyHostToDevice, s));
    if (shape.inner == 0) {
        EF_HIP(hipMemsetAsync(dc.ptr, 0, nc * 8, s));
    } else {
        EF_TRY(launch_dense_product_f64(da.as<double>(), db.as<double>(), dc.as<double>(), shape.batches,
                                        shape.rows, shape.columns, shape.inner, s));
    }
    EF_HIP(hipMemcpyAsync(out, dc.ptr, nc * 8, hipMemcpyDeviceToHost, s));
    EF_HIP(hipStreamSynchronize(s));
    return EF_OK;
}

ef_status ef_dense_run_attention(const double *q, const double *kt, const double *v, ef_attention_shape shape,
                                 uint32_t scale_kind, double scale, const size_t *allowed, double *out,
                                 const ef_options *options) {
    clear_error();
    hipStream_t s = as_stream(options ? options->stream : nullptr);
    const size_t B = shape.batches, X = shape.queries, E = shape.features, J = shape.keys, Y = shape.values;
    const size_t no = B * X * Y;
    if (no == 0) return EF_OK;
    // The kernels take keys [w][j][e]; the backend's are transposed [w][e][j].
    std::vector<double> k(B * J * E);
    for (size_t w = 0; w < B; w++)
        for (size_t e = 0; e < E; e++)
            for (size_t j = 0; j < J; j++) k[(w * J + j) * E + e] = kt[(w * E + e) * J + j];
    std::vector<int64_t> spans(2 * X);
    for (size_t x = 0; x < X; x++) {
        spans[2 * x] = (int64_t)allowed[2 * x];
        spans[2 * x + 1] = (int64_t)allowed[2 * x + 1];
    }
    DeviceBuffer dq, dk, dv, dout, dm, dl, dsp;
    EF_TRY(dq.alloc(B * X * E * 8));
    EF_TRY(dk.alloc(B * J * E * 8));
    EF_TRY(dv.alloc(B * J * Y * 8));
    EF_TRY(dout.alloc(no * 8));
    EF_TRY(dm.alloc(B * X * 8));
    EF_TRY(dl.alloc(B * X * 8));
    EF_TRY(dsp.alloc(2 * X * 8));
    if (B * X * E) EF_HIP(hipMemcpyAsync(dq.ptr, q, B * X * E * 8, hipMemcpyHostToDevice, s));
    if (B * J * E) EF_HIP(hipMemcpyAsync(dk.ptr, k.data(), B * J * E * 8, hipMemcpyHostToDevice, s));
    if (B * J * Y) EF_HIP(hipMemcpyAsync(dv.ptr, v, B * J * Y * 8, hipMemcpyHostToDevice, s));
    EF_HIP(hipMemcpyAsync(dsp.ptr, spans.data(), 2 * X * 8, hipMemcpyHostToDevice, s));
    ef_attention_desc d;
    std::memset(&d, 0, sizeof d);
    d.batches = B; d.queries = X; d.keys = J; d.features = E; d.values = Y;
    // Scale::Over(c) divides each score by c (backend.rs:49-86); the kernels
    // multiply by the rounded 1/c, one more rounding per score, which the
    // backend contract's tolerance absorbs (its bits are not pinned).
    d.scale = scale_kind == EF_SCALE_TIMES ? scale : (scale_kind == EF_SCALE_OVER ? 1.0 / scale : 1.0);
    d.allowed = dsp.as<int64_t>();
    d.precision = EF_PRECISION_FP64;
    DeviceBuffer scratch;
    EF_TRY(scratch.alloc(attention_scratch_bytes(d)));
    EF_TRY(attention_fwd(d, dq.ptr, dk.ptr, dv.ptr, dout.ptr, dm.ptr, dl.ptr, nullptr, scratch.ptr, s));
    EF_HIP(hipMemcpyAsync(out, dout.ptr, no * 8, hipMemcpyDeviceToHost, s));
    EF_HIP(hipStreamSynchronize(s));
    return EF_OK;
}

}  // extern "C"
