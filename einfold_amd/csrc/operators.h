// Host-side pieces shared by the operator entry points: the fold
// description read into plain vectors, the output schema, the budget and
// device allocations of one call, and uploads of Arrow columns.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <deque>
#include <string>
#include <utility>
#include <vector>

#include "arrow_io.h"
#include "common.h"

namespace ef {

// One operand of a fold (fold.rs:101-111).
struct OperandModel {
    std::vector<std::pair<uint32_t, uint32_t>> keys;   // (dimension, column)
    int factor = -1;                                   // -1: a mask
    int column_of(uint32_t d) const {
        for (auto &k : keys)
            if (k.first == d) return (int)k.second;
        return -1;
    }
};

// A fold (fold.rs:380-388) checked for indices in range.
struct FoldModel {
    std::vector<uint32_t> equality;                   // per dimension
    std::vector<OperandModel> operands;
    std::vector<ef_output> outputs;
    std::vector<ef_condition> conditions;
    bool has_domain = false;
    std::vector<uint32_t> carry;
    bool has_missing = false;
    double missing = 0.0;
    std::vector<ef_expr_op> row;                       // empty: the product
    int output_of(uint32_t d) const {
        for (size_t i = 0; i < outputs.size(); i++)
            if (outputs[i].dimension == d) return (int)i;
        return -1;
    }
    bool aggregated(uint32_t d) const { return output_of(d) < 0; }
};

bool parse_fold(const ef_fold_desc *f, FoldModel &m);

// schema.rs:40-143: output keys typed as the first operand holding each
// dimension, then a Float64 value, nullable unless a domain gives a
// default.  With a domain, its carried columns come first instead.
bool fold_schema(const FoldModel &m, const std::vector<const Relation *> &operands, const Relation *domain,
                 bool accept_float32, std::vector<OutField> &fields);

// Device memory chunks kept between calls (per device, up to a cap), so
// that calls in steady state allocate nothing: chunk_take gives the
// smallest cached chunk of at least `bytes` (false: none), chunk_give keeps
// or frees a chunk the device no longer uses, chunk_drop frees them all
// (before retrying an allocation that ran short).
bool chunk_take(size_t bytes, DeviceBuffer &out);
void chunk_give(DeviceBuffer &&chunk);
void chunk_drop();

// One call's options, stream, budget and device memory.  The call's
// device buffers are carved from a few large chunks (at least 4 MiB,
// doubling), taken from the chunk cache where it has one: the dozens of
// small buffers a fold needs cost no hipMalloc once a call of that size
// has run.  When the call ends its stream is synchronized and the chunks go
// back to the cache.
constexpr size_t CHUNK_GROW_MAX = (size_t)1 << 30;   // largest chunk carved for many buffers

struct Call {
    const ef_options *options = nullptr;
    hipStream_t stream = nullptr;
    std::deque<DeviceBuffer> buffers;   // the chunks
    size_t used = 0;                    // bytes carved from the last chunk
    bool zero_run = false;
    struct PendingZero { char *p; size_t bytes, chunk; };
    std::vector<PendingZero> pending_zeros;

    explicit Call(const ef_options *o) : options(o), stream(as_stream(o ? o->stream : nullptr)) {}
    ~Call() {
        if (buffers.empty()) return;
        (void)hipStreamSynchronize(stream);
        for (auto &b : buffers) chunk_give(std::move(b));
    }
    Call(const Call &) = delete;
    Call &operator=(const Call &) = delete;

    // options.rs:95-103: asks the budget, EF_OUT_OF_MEMORY if it says no.
    ef_status reserve(double bytes, const char *what);

    // n bytes of device memory, 256-byte aligned, alive until the call ends.
    ef_status carve(size_t n, void *&p) {
        n = (n + 255) & ~(size_t)255;
        if (buffers.empty() || used + n > buffers.back().bytes) {
            const size_t last = buffers.empty() ? 0 : buffers.back().bytes;
            // Chunks double from 4 MiB up to 1 GiB, so a call's many small
            // buffers take few allocations; a larger request gets a chunk of
            // its own size, so a call never holds much more than it uses.
            size_t chunk = (size_t)4 << 20;
            if (chunk < 2 * last) chunk = 2 * last;
            if (chunk > CHUNK_GROW_MAX) chunk = CHUNK_GROW_MAX;
            if (chunk < n) chunk = n;
            buffers.emplace_back();
            if (!chunk_take(n, buffers.back())) {
                ef_status st = buffers.back().alloc(chunk);
                if (st != EF_OK) {
                    chunk_drop();   // the cache's chunks back to the device, then just what is asked
                    st = buffers.back().alloc(n);
                }
                if (st != EF_OK) {
                    buffers.pop_back();
                    p = nullptr;
                    return st;
                }
            }
            used = 0;
        }
        p = static_cast<char *>(buffers.back().ptr) + used;
        used += n;
        return EF_OK;
    }

    template <class T> ef_status alloc(T *&p, size_t n) {
        void *v = nullptr;
        ef_status st = carve(n * sizeof(T) + 16, v);
        p = static_cast<T *>(v);
        return st;
    }
    // n zeroed T.  Between begin_zeros() and end_zeros() the buffers are
    // only carved, and end_zeros() clears them with one fill per chunk they
    // lie in (carved back to back there: the alignment gaps between them,
    // this call's too, are cleared with them).
    template <class T> ef_status zeros(T *&p, size_t n) {
        EF_TRY(alloc(p, n));
        if (zero_run) {
            pending_zeros.push_back({reinterpret_cast<char *>(p), n * sizeof(T) + 16, buffers.size()});
            return EF_OK;
        }
        EF_HIP(hipMemsetAsync(p, 0, n * sizeof(T) + 16, stream));
        return EF_OK;
    }
    void begin_zeros() {
        zero_run = true;
        pending_zeros.clear();
    }
    ef_status end_zeros() {
        zero_run = false;
        size_t i = 0;
        while (i < pending_zeros.size()) {
            char *lo = pending_zeros[i].p, *hi = lo + pending_zeros[i].bytes;
            size_t j = i + 1;
            for (; j < pending_zeros.size() && pending_zeros[j].chunk == pending_zeros[i].chunk; j++)
                hi = pending_zeros[j].p + pending_zeros[j].bytes;
            EF_HIP(hipMemsetAsync(lo, 0, (size_t)(hi - lo), stream));
            i = j;
        }
        pending_zeros.clear();
        return EF_OK;
    }
    // n T from host memory (pageable) to the device; the host buffer may
    // be reused when this returns.
    template <class T> ef_status upload(T *dst, const void *src, size_t n) {
        return upload_bytes(dst, src, n * sizeof(T));
    }
    ef_status upload_bytes(void *dst, const void *src, size_t bytes);
    template <class T> ef_status download(void *dst, const T *src, size_t n) {
        if (n) EF_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, stream));
        return EF_OK;
    }
    ef_status sync() {
        EF_HIP(hipStreamSynchronize(stream));
        return EF_OK;
    }
};

// A column of every batch of a relation, on the device as one array:
// factors as f64 with a validity byte per row, keys as i64.
ef_status upload_factor(Call &c, const Relation &r, int column, double *values, uint8_t *valid, bool *any_null);
ef_status upload_validity(Call &c, const Relation &r, int column, uint8_t *valid);
ef_status upload_raw(Call &c, const Relation &r, int column, void *dst);

// Host copy of one input column into an owned output column (a domain's
// carried columns).
OutColumn copy_column(const ColumnView &v, int64_t rows, bool nullable);

// Builds an output column from device values (width bytes each) and an
// optional device byte-per-row validity.  The copies are only enqueued on
// the call's stream: the column is complete once the stream has been
// synchronized and settle() has run.
ef_status fetch_column(Call &c, Type t, const void *values, const uint8_t *valid, int64_t n, bool nullable,
                       OutColumn &out);
// Completes the columns fetch_column built (their NULL counts); the
// stream that copied them must have been synchronized.
void settle(std::vector<std::vector<OutColumn>> &batches);

// Moves batches and a schema into an ef_result (settling them first: their
// copies must have finished).
void put_result(std::vector<std::vector<OutColumn>> batches, const std::vector<OutField> &fields, uint32_t algorithm,
                ef_result *out);

}  // namespace ef
