// Arrow C Data Interface on the host: typed views of a relation's columns,
// and owned output batches whose buffers are released through the
// interface's release callbacks.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/einfold_amd.h"

namespace ef {

enum class Type { Int8, UInt8, Int16, UInt16, Int32, UInt32, Int64, UInt64, Float32, Float64, Other };

Type type_of(const char *format);
int width_of(Type t);              // bytes per value, 0 for Other
bool is_integer(Type t);
bool is_float(Type t);
bool is_signed(Type t);
const char *format_of(Type t);

// One column of one batch: its values (already offset), validity bitmap
// with its bit offset (or null), length and null count.
struct ColumnView {
    Type type = Type::Other;
    const uint8_t *values = nullptr;   // element 0 of this batch's rows
    const uint8_t *validity = nullptr; // bitmap, or null when no NULLs
    int64_t bit_offset = 0;            // of element 0 in the bitmap
    int64_t length = 0;
    int64_t null_count = 0;
    bool valid(int64_t r) const {
        if (!validity) return true;
        const int64_t b = bit_offset + r;
        return (validity[b >> 3] >> (b & 7)) & 1;
    }
};

// A relation: its field types and each batch's columns.
struct Relation {
    std::vector<Type> types;
    std::vector<bool> nullable;
    std::vector<std::string> formats;
    std::vector<std::vector<ColumnView>> batches;   // [batch][column]
    std::vector<int64_t> batch_rows;
    int64_t rows = 0;
    size_t columns() const { return types.size(); }
};

// Reads an ef_relation; false (with the error set) if malformed.
bool read_relation(const ef_relation *in, Relation &out);

// Host buffer aligned to 64 bytes, as Arrow recommends; pinned (from a
// pool of released pinned buffers) from 64 KiB.
struct HostBuffer {
    uint8_t *data = nullptr;
    size_t bytes = 0;
    size_t capacity = 0;   // a pinned buffer's size class
    bool pinned = false;
    explicit HostBuffer(size_t n);
    ~HostBuffer();
    HostBuffer(const HostBuffer &) = delete;
    HostBuffer &operator=(const HostBuffer &) = delete;
};

// An output column under construction: type, values, optional validity.
struct OutColumn {
    Type type = Type::Float64;
    bool nullable = true;
    int64_t length = 0;
    int64_t null_count = 0;
    std::shared_ptr<HostBuffer> values;
    std::shared_ptr<HostBuffer> validity;   // null: no NULLs
    // Set while the column's NULL count is still being copied from the
    // device: after the stream is synchronized, settle() reads it and drops
    // a bitmap that marks no NULL.
    std::shared_ptr<unsigned long long> pending_nulls;
};

struct OutField {
    Type type;
    bool nullable;
};

// Moves the columns into an Arrow struct array (a record batch).
void export_batch(std::vector<OutColumn> columns, struct ArrowArray *out);
// An Arrow struct schema with fields named "0", "1", ... .
void export_schema(const std::vector<OutField> &fields, struct ArrowSchema *out);

}  // namespace ef
