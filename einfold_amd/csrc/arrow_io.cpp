// Arrow C Data Interface import and export (see arrow_io.h).
#include "arrow_io.h"

#include <map>
#include <mutex>

#include <cstdlib>
#include <cstring>
#include <unistd.h>

#include <hip/hip_runtime.h>
#include <new>

#include "common.h"

namespace ef {

Type type_of(const char *f) {
    if (!f || !f[0] || f[1]) return Type::Other;
    switch (f[0]) {
    case 'c': return Type::Int8;
    case 'C': return Type::UInt8;
    case 's': return Type::Int16;
    case 'S': return Type::UInt16;
    case 'i': return Type::Int32;
    case 'I': return Type::UInt32;
    case 'l': return Type::Int64;
    case 'L': return Type::UInt64;
    case 'f': return Type::Float32;
    case 'g': return Type::Float64;
    default: return Type::Other;
    }
}

int width_of(Type t) {
    switch (t) {
    case Type::Int8: case Type::UInt8: return 1;
    case Type::Int16: case Type::UInt16: return 2;
    case Type::Int32: case Type::UInt32: case Type::Float32: return 4;
    case Type::Int64: case Type::UInt64: case Type::Float64: return 8;
    default: return 0;
    }
}

bool is_integer(Type t) { return t <= Type::UInt64; }
bool is_float(Type t) { return t == Type::Float32 || t == Type::Float64; }
bool is_signed(Type t) {
    return t == Type::Int8 || t == Type::Int16 || t == Type::Int32 || t == Type::Int64;
}

const char *format_of(Type t) {
    static const char *names[] = {"c", "C", "s", "S", "i", "I", "l", "L", "f", "g", ""};
    return names[(int)t];
}

static int64_t count_nulls(const uint8_t *bitmap, int64_t offset, int64_t n) {
    if (!bitmap) return 0;
    int64_t nulls = 0;
    for (int64_t r = 0; r < n; r++) {
        const int64_t b = offset + r;
        nulls += ((bitmap[b >> 3] >> (b & 7)) & 1) ? 0 : 1;
    }
    return nulls;
}

bool read_relation(const ef_relation *in, Relation &out) {
    out = Relation();
    if (!in || !in->schema) {
        set_error("a relation needs a schema");
        return false;
    }
    const ArrowSchema *s = in->schema;
    if (!s->format || std::strcmp(s->format, "+s") != 0) {
        set_error("a relation's schema must be a struct (record batch) schema");
        return false;
    }
    for (int64_t c = 0; c < s->n_children; c++) {
        const ArrowSchema *f = s->children[c];
        out.types.push_back(type_of(f->format));
        out.formats.push_back(f->format ? f->format : "");
        out.nullable.push_back((f->flags & ARROW_FLAG_NULLABLE) != 0);
    }
    for (size_t b = 0; b < in->n_batches; b++) {
        const ArrowArray *a = in->batches[b];
        if (!a || a->n_children != s->n_children) {
            set_error("a batch doesn't match its relation's schema");
            return false;
        }
        std::vector<ColumnView> cols(s->n_children);
        for (int64_t c = 0; c < s->n_children; c++) {
            const ArrowArray *ch = a->children[c];
            ColumnView &v = cols[c];
            v.type = out.types[c];
            v.length = a->length;
            if (ch->length < a->offset + a->length) {
                set_error("a column is shorter than its batch");
                return false;
            }
            const int64_t off = ch->offset + a->offset;
            const int w = width_of(v.type);
            if (w > 0 && ch->n_buffers >= 2) {
                v.values = static_cast<const uint8_t *>(ch->buffers[1]);
                if (v.values) v.values += off * w;
                v.validity = static_cast<const uint8_t *>(ch->buffers[0]);
                v.bit_offset = off;
                if (v.validity) {
                    if (ch->null_count == 0 && a->offset == 0) v.validity = nullptr;
                    else v.null_count = count_nulls(v.validity, off, v.length);
                    if (v.null_count == 0) v.validity = nullptr;
                }
            }
        }
        out.batches.push_back(std::move(cols));
        out.batch_rows.push_back(a->length);
        out.rows += a->length;
    }
    return true;
}

// Output buffers from 64 KiB up are pinned: the results' device-to-host
// copies then run as direct DMA into them, where a pageable buffer would be
// staged in chunks (and, fresh, take a page fault on first touch of each
// page: about 13x slower for hundreds of MB).  Pinning a buffer
// (hipHostMalloc) costs far more than copying into it, so released pinned
// buffers are kept in a pool by size class and handed out again, up to
// pin_pool_cap() bytes (a sixteenth of the host's memory, at most 16 GiB:
// room for the outputs of one call of a few hundred million rows); the
// pool lives until the process ends (its buffers are never freed after the
// runtime has shut down).  Smaller buffers stay pageable.
namespace {
constexpr size_t PIN_FROM = (size_t)64 << 10;

size_t pin_pool_cap() {
    static const size_t cap = [] {
        const long pages = sysconf(_SC_PHYS_PAGES), page = sysconf(_SC_PAGESIZE);
        const size_t host = pages > 0 && page > 0 ? (size_t)pages * (size_t)page : 0;
        const size_t sixteenth = host / 16;
        const size_t most = (size_t)16 << 30, least = (size_t)1 << 30;
        return sixteenth > most ? most : (sixteenth < least ? least : sixteenth);
    }();
    return cap;
}

struct PinnedPool {
    std::mutex mu;
    std::multimap<size_t, void *> free;   // size class -> buffer
    size_t cached = 0;
};

PinnedPool &pinned_pool() {
    static PinnedPool *pool = new PinnedPool;
    return *pool;
}

// Size classes: powers of two up to 1 MiB, then multiples of an eighth of
// the power of two below (at most 12.5 % over the size asked).
size_t size_class(size_t n) {
    size_t c = PIN_FROM;
    while (c < n) c <<= 1;
    if (c <= ((size_t)1 << 20)) return c;
    const size_t step = (c >> 1) / 8;
    return (n + step - 1) / step * step;
}
}  // namespace

HostBuffer::HostBuffer(size_t n) : bytes(n) {
    const size_t padded = ((n + 63) / 64) * 64 + 64;
    if (padded >= PIN_FROM) {
        const size_t cls = size_class(padded);
        PinnedPool &pool = pinned_pool();
        {
            std::lock_guard<std::mutex> lock(pool.mu);
            auto it = pool.free.find(cls);
            if (it != pool.free.end()) {
                data = static_cast<uint8_t *>(it->second);
                pool.cached -= cls;
                pool.free.erase(it);
            }
        }
        if (!data) {
            void *p = nullptr;
            if (hipHostMalloc(&p, cls, hipHostMallocDefault) == hipSuccess && p) data = static_cast<uint8_t *>(p);
            else (void)hipGetLastError();
        }
        if (data) {
            pinned = true;
            capacity = cls;
        }
    }
    if (!data) data = static_cast<uint8_t *>(std::aligned_alloc(64, padded));
    if (!data) throw std::bad_alloc();
    std::memset(data + (padded - 64), 0, 64);
}

HostBuffer::~HostBuffer() {
    if (!pinned) {
        std::free(data);
        return;
    }
    PinnedPool &pool = pinned_pool();
    {
        std::lock_guard<std::mutex> lock(pool.mu);
        if (pool.cached + capacity <= pin_pool_cap()) {
            pool.free.emplace(capacity, data);
            pool.cached += capacity;
            return;
        }
    }
    (void)hipHostFree(data);
}

namespace {

struct ArrayPrivate {
    std::vector<std::shared_ptr<HostBuffer>> owned;
    std::vector<const void *> buffers;
    std::vector<ArrowArray *> children;
};

void release_array(ArrowArray *a) {
    if (!a || !a->release) return;
    for (int64_t i = 0; i < a->n_children; i++) {
        ArrowArray *ch = a->children[i];
        if (ch->release) ch->release(ch);
        delete ch;
    }
    delete static_cast<ArrayPrivate *>(a->private_data);
    a->release = nullptr;
}

struct SchemaPrivate {
    std::string format, name;
    std::vector<ArrowSchema *> children;
};

void release_schema(ArrowSchema *s) {
    if (!s || !s->release) return;
    for (int64_t i = 0; i < s->n_children; i++) {
        ArrowSchema *ch = s->children[i];
        if (ch->release) ch->release(ch);
        delete ch;
    }
    delete static_cast<SchemaPrivate *>(s->private_data);
    s->release = nullptr;
}

}  // namespace

void export_batch(std::vector<OutColumn> columns, ArrowArray *out) {
    auto *top = new ArrayPrivate();
    int64_t length = columns.empty() ? 0 : columns[0].length;
    for (OutColumn &col : columns) {
        auto *cp = new ArrayPrivate();
        cp->owned.push_back(col.values);
        if (col.validity) cp->owned.push_back(col.validity);
        cp->buffers = {col.validity ? col.validity->data : nullptr, col.values ? col.values->data : nullptr};
        auto *ch = new ArrowArray();
        ch->length = col.length;
        ch->null_count = col.null_count;
        ch->offset = 0;
        ch->n_buffers = 2;
        ch->n_children = 0;
        ch->buffers = cp->buffers.data();
        ch->children = nullptr;
        ch->dictionary = nullptr;
        ch->release = release_array;
        ch->private_data = cp;
        top->children.push_back(ch);
    }
    top->buffers = {nullptr};
    out->length = length;
    out->null_count = 0;
    out->offset = 0;
    out->n_buffers = 1;
    out->n_children = (int64_t)top->children.size();
    out->buffers = top->buffers.data();
    out->children = top->children.data();
    out->dictionary = nullptr;
    out->release = release_array;
    out->private_data = top;
}

void export_schema(const std::vector<OutField> &fields, ArrowSchema *out) {
    auto *top = new SchemaPrivate();
    top->format = "+s";
    for (size_t i = 0; i < fields.size(); i++) {
        auto *cp = new SchemaPrivate();
        cp->format = format_of(fields[i].type);
        cp->name = std::to_string(i);
        auto *ch = new ArrowSchema();
        ch->format = cp->format.c_str();
        ch->name = cp->name.c_str();
        ch->metadata = nullptr;
        ch->flags = fields[i].nullable ? ARROW_FLAG_NULLABLE : 0;
        ch->n_children = 0;
        ch->children = nullptr;
        ch->dictionary = nullptr;
        ch->release = release_schema;
        ch->private_data = cp;
        top->children.push_back(ch);
    }
    out->format = top->format.c_str();
    out->name = "";
    out->metadata = nullptr;
    out->flags = 0;
    out->n_children = (int64_t)top->children.size();
    out->children = top->children.data();
    out->dictionary = nullptr;
    out->release = release_schema;
    out->private_data = top;
}

}  // namespace ef
