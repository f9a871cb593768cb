// The kernel API that EinFoldExec calls (crates/einfold-kernels/src/
// algorithm.rs:52-160): fold descriptions and Arrow batches in, Arrow
// batches out, with the device doing the per-row work.  This file holds the
// shared host pieces and the dense positional fold (dense.rs:176-650).
#include "operators.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <condition_variable>
#include <thread>
#include <functional>
#include <limits>
#include <map>
#include <mutex>

#include "discriminate.h"
#include "pipeline.h"
#include "ops.h"

namespace ef {

namespace {

constexpr uint64_t U32_MAX = 0xffffffffULL;
constexpr uint64_t FEW_CELLS = 4096;          // dense.rs:85-91
constexpr uint64_t SATURATED = ~0ULL;

// Saturating products of extents (a size past 2^64 fails every check).
uint64_t mul_sat(uint64_t a, uint64_t b) {
    const unsigned __int128 p = (unsigned __int128)a * b;
    return p > SATURATED ? SATURATED : (uint64_t)p;
}

uint64_t size_of(const std::vector<uint32_t> &dims, const std::vector<uint64_t> &extent) {
    uint64_t s = 1;
    for (uint32_t d : dims) s = mul_sat(s, extent[d]);
    return s;
}

// dense.rs:85-91: a matrix is worth it with at most four cells per input
// row (plus slack for small inputs), and positions must fit 32 bits.
bool worth(uint64_t cells, uint64_t rows) {
    return cells <= mul_sat(rows, 4) + FEW_CELLS && cells < U32_MAX;
}

int index_in(const std::vector<uint32_t> &v, uint32_t d) {
    for (size_t i = 0; i < v.size(); i++)
        if (v[i] == d) return (int)i;
    return -1;
}

// The stride of dimension d within its role (later dimensions vary faster).
uint64_t stride_in(const std::vector<uint32_t> &role, uint32_t d, const std::vector<uint64_t> &extent) {
    uint64_t s = 1;
    for (size_t i = (size_t)index_in(role, d) + 1; i < role.size(); i++) s *= extent[role[i]];
    return s;
}

bool factor_type_ok(Type t, bool accept_float32) {
    return t == Type::Float64 || (accept_float32 && t == Type::Float32);
}

}  // namespace

// ---------------------------------------------------------------------
// Shared pieces
// ---------------------------------------------------------------------

bool parse_fold(const ef_fold_desc *f, FoldModel &m) {
    m = FoldModel();
    if (!f) {
        set_error("no fold description");
        return false;
    }
    const uint32_t nd = f->n_dimensions;
    if (nd && !f->equality) {
        set_error("a fold's dimensions need their key equalities");
        return false;
    }
    m.equality.assign(f->equality, f->equality + nd);
    for (uint32_t e : m.equality)
        if (e > EF_NOT_DISTINCT) {
            set_error("unknown key equality");
            return false;
        }
    if (f->n_operands && !f->operands) {
        set_error("operands missing");
        return false;
    }
    for (uint32_t o = 0; o < f->n_operands; o++) {
        const ef_operand &op = f->operands[o];
        OperandModel om;
        om.factor = op.factor;
        for (uint32_t k = 0; k < op.n_keys; k++) {
            if (op.keys[k].dimension >= nd) {
                set_error("a key names a dimension the fold doesn't have");
                return false;
            }
            om.keys.emplace_back(op.keys[k].dimension, op.keys[k].column);
        }
        m.operands.push_back(om);
    }
    for (uint32_t i = 0; i < f->n_outputs; i++) {
        if (f->outputs[i].dimension >= nd) {
            set_error("an output names a dimension the fold doesn't have");
            return false;
        }
        m.outputs.push_back(f->outputs[i]);
    }
    for (uint32_t i = 0; i < f->n_conditions; i++) {
        const ef_condition &c = f->conditions[i];
        if (c.left >= nd || c.right >= nd || c.comparison > EF_GREATER_OR_EQUAL) {
            set_error("a condition names a dimension the fold doesn't have");
            return false;
        }
        m.conditions.push_back(c);
    }
    if (f->domain) {
        m.has_domain = true;
        m.carry.assign(f->domain->carry, f->domain->carry + f->domain->n_carry);
        m.has_missing = f->domain->has_missing != 0;
        m.missing = f->domain->missing;
        for (auto &o : m.outputs)
            if (o.domain_column < 0) {
                set_error("with a domain every output needs its domain column");
                return false;
            }
    }
    if (f->row && f->n_row) m.row.assign(f->row, f->row + f->n_row);
    return true;
}

bool fold_schema(const FoldModel &m, const std::vector<const Relation *> &operands, const Relation *domain,
                 bool accept_float32, std::vector<OutField> &fields) {
    fields.clear();
    if (m.operands.size() != operands.size()) {
        set_error("one relation per operand");
        return false;
    }
    if (m.has_domain != (domain != nullptr)) {
        set_error("a domain relation must be given exactly when the fold has a domain");
        return false;
    }
    const size_t nd = m.equality.size();
    std::vector<int> have(nd, 0);
    std::vector<OutField> dim(nd, OutField{Type::Other, true});
    for (size_t o = 0; o < m.operands.size(); o++) {
        const OperandModel &op = m.operands[o];
        const Relation &r = *operands[o];
        if (m.row.empty()) {
            if (op.factor < 0) {
                set_error("EinFold doesn't compute masks yet");
                return false;
            }
            if ((size_t)op.factor >= r.columns() || !factor_type_ok(r.types[op.factor], accept_float32)) {
                set_error("factors must be Float64");
                return false;
            }
        }
        for (auto &k : op.keys) {
            if (k.second >= r.columns()) {
                set_error("a key's column is past its relation's columns");
                return false;
            }
            const OutField f{r.types[k.second], (bool)r.nullable[k.second]};
            if (!have[k.first]) {
                have[k.first] = 1;
                dim[k.first] = f;
            } else if (dim[k.first].type != f.type) {
                set_error("dimension " + std::to_string(k.first) + "'s columns have two types");
                return false;
            }
        }
    }
    for (const ef_expr_op &e : m.row) {
        if (e.code != EF_EXPR_COLUMN) continue;
        if (e.operand >= operands.size() || e.column >= operands[e.operand]->columns() ||
            !factor_type_ok(operands[e.operand]->types[e.column], accept_float32)) {
            set_error("an expression's columns must be Float64");
            return false;
        }
    }
    if (!m.has_domain) {
        for (auto &o : m.outputs) {
            if (!have[o.dimension]) {
                set_error("an output's dimension is in no operand");
                return false;
            }
            fields.push_back(dim[o.dimension]);
        }
        fields.push_back(OutField{Type::Float64, true});
    } else {
        for (uint32_t c : m.carry) {
            if (c >= domain->columns()) {
                set_error("a carried column is past the domain's columns");
                return false;
            }
            fields.push_back(OutField{domain->types[c], (bool)domain->nullable[c]});
        }
        fields.push_back(OutField{Type::Float64, !m.has_missing});
    }
    return true;
}

// Device chunks kept between calls, up to a quarter of the device's memory
// in all: allocating device memory that was freed shortly before can cost
// a second or more per GiB (the driver clears it first), so a workload
// that repeats large calls must find its chunks here rather than free and
// allocate them again.  A call that runs short of memory returns the
// cache's chunks to the device first (Call::carve).
namespace {
std::mutex chunk_mu;
std::multimap<std::pair<int, size_t>, void *> chunk_cache;   // (device, bytes) -> chunk
size_t chunk_cached = 0;

size_t chunk_cap() {
    static size_t cap = 0;
    if (!cap) {
        size_t free_b = 0, total_b = 0;
        cap = hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b ? total_b / 4 : (size_t)2 << 30;
    }
    return cap;
}
}  // namespace

bool chunk_take(size_t bytes, DeviceBuffer &out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    std::lock_guard<std::mutex> lock(chunk_mu);
    auto it = chunk_cache.lower_bound({dev, bytes});
    if (it == chunk_cache.end() || it->first.first != dev) return false;
    out.ptr = it->second;
    out.bytes = it->first.second;
    chunk_cached -= out.bytes;
    chunk_cache.erase(it);
    return true;
}

void chunk_give(DeviceBuffer &&chunk) {
    if (!chunk.ptr) return;
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
        std::lock_guard<std::mutex> lock(chunk_mu);
        if (chunk_cached + chunk.bytes <= chunk_cap()) {
            chunk_cache.emplace(std::make_pair(dev, chunk.bytes), chunk.ptr);
            chunk_cached += chunk.bytes;
            chunk.ptr = nullptr;
            chunk.bytes = 0;
            return;
        }
    }
    chunk.alloc(0);   // frees it
}

void chunk_drop() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    std::lock_guard<std::mutex> lock(chunk_mu);
    for (auto it = chunk_cache.begin(); it != chunk_cache.end();) {
        if (it->first.first != dev) {
            ++it;
            continue;
        }
        (void)hipFree(it->second);
        chunk_cached -= it->first.second;
        it = chunk_cache.erase(it);
    }
}

// A small pool of host threads for the library's own host work: run(f)
// calls f(0) .. f(POOL_THREADS - 1) at once, f(0) on the caller, and
// returns when all are done; one job at a time.  The workers live until the
// process ends.
namespace {
constexpr int POOL_THREADS = 4;

struct WorkPool {
    std::mutex run_mu;   // one job at a time
    std::mutex qmu;
    std::condition_variable go, done;
    uint64_t generation = 0;
    int finished = 0;
    const std::function<void(int)> *job = nullptr;
    bool started = false;

    void worker(int q) {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(int)> *f;
            {
                std::unique_lock<std::mutex> lock(qmu);
                go.wait(lock, [&] { return generation != seen; });
                seen = generation;
                f = job;
            }
            (*f)(q);
            {
                std::lock_guard<std::mutex> lock(qmu);
                finished++;
            }
            done.notify_one();
        }
    }
    void run(const std::function<void(int)> &f) {
        std::lock_guard<std::mutex> once(run_mu);
        if (!started) {
            for (int q = 1; q < POOL_THREADS; q++) std::thread([this, q] { worker(q); }).detach();
            started = true;
        }
        {
            std::lock_guard<std::mutex> lock(qmu);
            job = &f;
            finished = 0;
            generation++;
        }
        go.notify_all();
        f(0);
        std::unique_lock<std::mutex> lock(qmu);
        done.wait(lock, [&] { return finished == POOL_THREADS - 1; });
    }
};

WorkPool &work_pool() {
    static WorkPool *p = new WorkPool;   // lives until the process ends, its workers too
    return *p;
}
}  // namespace

// Large host-to-device copies from pageable memory: HIP stages those
// through its own pinned buffers with one host thread (about 30 GB/s
// here).  Copied instead by the work pool's threads into a ring of pinned
// chunks, each sent by DMA as soon as it is full while the next fills,
// they run at close to the link's 57 GB/s.  One staged copy at a time in
// the process; a chunk is refilled only after its DMA has completed (its
// event), and every byte is in pinned memory before upload_bytes returns,
// so the caller's buffer may be reused at once.
namespace {
constexpr size_t STAGE_FROM = (size_t)4 << 20;
constexpr size_t STAGE_CHUNK = (size_t)16 << 20;
constexpr int STAGE_SLOTS = 4;

struct Stager {
    std::mutex mu;   // one staged copy at a time
    char *ring = nullptr;
    hipEvent_t ev[STAGE_SLOTS] = {};
    bool pending[STAGE_SLOTS] = {};
    int next = 0;   // the slot the next chunk takes: successive copies rotate through the ring
    bool ready = false, failed = false;

    // Whether the ring exists (made at the first staged copy).
    bool start() {
        if (ready || failed) return ready;
        void *p = nullptr;
        if (hipHostMalloc(&p, STAGE_CHUNK * STAGE_SLOTS, hipHostMallocDefault) != hipSuccess || !p) {
            (void)hipGetLastError();
            failed = true;
            return false;
        }
        ring = static_cast<char *>(p);
        for (int i = 0; i < STAGE_SLOTS; i++) {
            if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) {
                (void)hipGetLastError();
                failed = true;
                return false;
            }
        }
        ready = true;
        return true;
    }
    static void copy(const char *s, char *d, size_t n) {
        const std::function<void(int)> f = [=](int q) {
            const size_t per = (n + POOL_THREADS - 1) / POOL_THREADS;
            const size_t a = per * (size_t)q, b = a + per < n ? a + per : n;
            if (a < b) std::memcpy(d + a, s + a, b - a);
        };
        work_pool().run(f);
    }
};

Stager &stager() {
    static Stager *s = new Stager;   // lives until the process ends
    return *s;
}
}  // namespace

ef_status Call::upload_bytes(void *dst, const void *src, size_t bytes) {
    if (!bytes) return EF_OK;
    if (bytes >= STAGE_FROM) {
        Stager &st = stager();
        std::lock_guard<std::mutex> lock(st.mu);
        if (st.start()) {
            for (size_t off = 0; off < bytes; off += STAGE_CHUNK) {
                const int k = st.next;
                st.next = (k + 1) % STAGE_SLOTS;
                const size_t n = bytes - off < STAGE_CHUNK ? bytes - off : STAGE_CHUNK;
                char *slot = st.ring + (size_t)k * STAGE_CHUNK;
                if (st.pending[k]) {
                    EF_HIP(hipEventSynchronize(st.ev[k]));
                    st.pending[k] = false;
                }
                st.copy(static_cast<const char *>(src) + off, slot, n);
                EF_HIP(hipMemcpyAsync(static_cast<char *>(dst) + off, slot, n, hipMemcpyHostToDevice, stream));
                EF_HIP(hipEventRecord(st.ev[k], stream));
                st.pending[k] = true;
            }
            return EF_OK;
        }
    }
    EF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
    return EF_OK;
}

ef_status Call::reserve(double bytes, const char *what) {
    if (!options || !options->budget_grow) return EF_OK;
    const size_t b = bytes >= 1.8e19 ? std::numeric_limits<size_t>::max() : (size_t)bytes;
    if (options->budget_grow(options->budget_ctx, b)) return EF_OK;
    set_error("EinFold's budget can't give " + std::to_string(b) + " bytes for " + what);
    return EF_OUT_OF_MEMORY;
}

namespace {

// Uploads one batch's bitmap bits [offset, offset + n) and unpacks them.
ef_status unpack_validity(Call &c, const ColumnView &v, uint8_t *dst, uint8_t *&stage, size_t &stage_bytes) {
    const int64_t first = v.bit_offset >> 3;
    const size_t bytes = (size_t)(((v.bit_offset + v.length + 7) >> 3) - first);
    if (bytes > stage_bytes) {
        EF_TRY(c.alloc(stage, bytes));
        stage_bytes = bytes;
    }
    EF_TRY(c.upload<uint8_t>(stage, v.validity + first, bytes));
    return unpack_bits(v.length, stage, v.bit_offset & 7, dst, c.stream);
}

}  // namespace

ef_status upload_validity(Call &c, const Relation &r, int column, uint8_t *valid) {
    uint8_t *stage = nullptr;
    size_t stage_bytes = 0;
    int64_t off = 0;
    for (size_t b = 0; b < r.batches.size(); b++) {
        const ColumnView &v = r.batches[b][column];
        if (v.validity) EF_TRY(unpack_validity(c, v, valid + off, stage, stage_bytes));
        else if (v.length) EF_HIP(hipMemsetAsync(valid + off, 1, v.length, c.stream));
        off += v.length;
    }
    return EF_OK;
}

ef_status upload_factor(Call &c, const Relation &r, int column, double *values, uint8_t *valid, bool *any_null) {
    float *stage32 = nullptr;
    size_t stage32_n = 0;
    int64_t off = 0;
    for (size_t b = 0; b < r.batches.size(); b++) {
        const ColumnView &v = r.batches[b][column];
        if (any_null && v.null_count) *any_null = true;
        if (v.type == Type::Float64) {
            EF_TRY(c.upload<double>(values + off, v.values, v.length));
        } else if (v.type == Type::Float32) {
            if ((size_t)v.length > stage32_n) {
                EF_TRY(c.alloc(stage32, v.length));
                stage32_n = v.length;
            }
            EF_TRY(c.upload<float>(stage32, v.values, v.length));
            EF_TRY(to_f64(Type::Float32, stage32, v.length, values + off, c.stream));
        } else {
            set_error("factors must be floats");
            return EF_INVALID;
        }
        off += v.length;
    }
    return valid ? upload_validity(c, r, column, valid) : EF_OK;
}

ef_status upload_raw(Call &c, const Relation &r, int column, void *dst) {
    const int w = width_of(r.types[column]);
    int64_t off = 0;
    for (size_t b = 0; b < r.batches.size(); b++) {
        const ColumnView &v = r.batches[b][column];
        EF_TRY(c.upload<uint8_t>(static_cast<uint8_t *>(dst) + off * w, v.values, (size_t)v.length * w));
        off += v.length;
    }
    return EF_OK;
}

OutColumn copy_column(const ColumnView &v, int64_t rows, bool nullable) {
    OutColumn o;
    o.type = v.type;
    o.nullable = nullable;
    o.length = rows;
    const int w = width_of(v.type);
    o.values = std::make_shared<HostBuffer>((size_t)rows * w);
    if (rows) std::memcpy(o.values->data, v.values, (size_t)rows * w);
    if (v.validity && v.null_count) {
        o.validity = std::make_shared<HostBuffer>((size_t)(rows + 7) / 8);
        std::memset(o.validity->data, 0, (size_t)(rows + 7) / 8);
        for (int64_t r = 0; r < rows; r++)
            if (v.valid(r)) o.validity->data[r >> 3] |= (uint8_t)(1u << (r & 7));
        o.null_count = v.null_count;
    }
    return o;
}

ef_status fetch_column(Call &c, Type t, const void *values, const uint8_t *valid, int64_t n, bool nullable,
                       OutColumn &out) {
    out = OutColumn();
    out.type = t;
    out.nullable = nullable;
    out.length = n;
    const int w = width_of(t);
    out.values = std::make_shared<HostBuffer>((size_t)n * w);
    EF_TRY(c.download<uint8_t>(out.values->data, static_cast<const uint8_t *>(values), (size_t)n * w));
    if (valid && n) {
        uint8_t *dbits;
        unsigned long long *dnulls;
        EF_TRY(c.alloc(dbits, (size_t)(n + 7) / 8));
        EF_TRY(c.zeros(dnulls, 1));
        EF_TRY(pack_bits(n, valid, dbits, dnulls, c.stream));
        out.validity = std::make_shared<HostBuffer>((size_t)(n + 7) / 8);
        out.pending_nulls = std::make_shared<unsigned long long>(0);
        EF_TRY(c.download<uint8_t>(out.validity->data, dbits, (size_t)(n + 7) / 8));
        EF_TRY(c.download<unsigned long long>(out.pending_nulls.get(), dnulls, 1));
    }
    return EF_OK;
}

void settle(std::vector<std::vector<OutColumn>> &batches) {
    for (auto &cols : batches)
        for (auto &col : cols) {
            if (!col.pending_nulls) continue;
            col.null_count = (int64_t)*col.pending_nulls;
            if (col.null_count == 0) col.validity.reset();
            col.pending_nulls.reset();
        }
}

void put_result(std::vector<std::vector<OutColumn>> batches, const std::vector<OutField> &fields, uint32_t algorithm,
                ef_result *out) {
    settle(batches);
    std::memset(out, 0, sizeof *out);
    export_schema(fields, &out->schema);
    out->n_batches = batches.size();
    out->batches = batches.empty() ? nullptr : new ArrowArray[batches.size()];
    for (size_t i = 0; i < batches.size(); i++) export_batch(std::move(batches[i]), &out->batches[i]);
    out->algorithm = algorithm;
}

// ---------------------------------------------------------------------
// The dense positional fold (dense.rs:176-578)
// ---------------------------------------------------------------------

// A batch of dense folds whose MFMA contractions run as one grouped launch:
// each fold appends its products and leaves `finish` (the widening of its
// sums and its emission) to run after that launch.  A fold that leaves
// `finish` empty has finished on its own.
struct DenseBatch {
    std::vector<ef_gemm_problem> problems;
    std::vector<double> density;
    uint32_t precision = EF_PRECISION_FP32;
    std::function<ef_status()> finish;
};

ef_status fold_dense_impl(const FoldModel &m, const Relation &a, const Relation &b, const Relation *domain,
                          const std::vector<OutField> &fields, Call &c,
                          std::vector<std::vector<OutColumn>> &result, DenseBatch *batch) {
    const bool accept32 = c.options && c.options->accept_float32;
    if (m.operands.size() != 2) return EF_DECLINED;
    const OperandModel &A = m.operands[0], &B = m.operands[1];
    const uint32_t nd = (uint32_t)m.equality.size();
    // Roles (dense.rs:196-230): W in both operands and the output, X in a
    // only, Y in b only (output order); Z aggregated, in both (dimension
    // order).  A row of the output is W's then X's.
    std::vector<uint32_t> ws, xs, ys, zs;
    for (auto &o : m.outputs) {
        const bool ina = A.column_of(o.dimension) >= 0, inb = B.column_of(o.dimension) >= 0;
        if (ina && inb) ws.push_back(o.dimension);
        else if (ina) xs.push_back(o.dimension);
        else if (inb) ys.push_back(o.dimension);
        else return EF_DECLINED;
    }
    xs.insert(xs.begin(), ws.begin(), ws.end());
    for (uint32_t d = 0; d < nd; d++) {
        if (!m.aggregated(d)) continue;
        if (A.column_of(d) < 0 || B.column_of(d) < 0) return EF_DECLINED;
        zs.push_back(d);
    }
    auto is_out = [&](uint32_t d) { return index_in(xs, d) >= 0 || index_in(ys, d) >= 0; };
    for (auto &cd : m.conditions)
        if (!is_out(cd.left) || !is_out(cd.right)) return EF_DECLINED;
    auto integers = [](const OperandModel &op, const Relation &r) {
        if (r.batches.empty()) return true;
        for (auto &k : op.keys)
            if (!is_integer(r.types[k.second])) return false;
        return true;
    };
    if (!integers(A, a) || !integers(B, b)) return EF_DECLINED;
    EF_TRY(c.reserve(((double)a.rows * A.keys.size() + (double)b.rows * B.keys.size()) * 8,
                     "the dense algorithm's keys"));
    // dense.rs:97-111: keys without NULLs, non-negative, below 2^63 - 1.
    for (auto *pr : {&a, &b}) {
        const OperandModel &op = pr == &a ? A : B;
        for (auto &bt : pr->batches)
            for (auto &k : op.keys)
                if (bt[k.second].null_count) return EF_DECLINED;
    }
    KeyStats *stats;
    EF_TRY(c.alloc(stats, nd ? nd : 1));
    {
        std::vector<KeyStats> init(nd ? nd : 1, KeyStats{0, -1, 0, 0});
        EF_TRY(c.upload<KeyStats>(stats, init.data(), init.size()));
    }
    // Each operand's keys as i64 per dimension, every batch in one array.
    std::vector<int64_t *> a_key(nd, nullptr), b_key(nd, nullptr);
    size_t raw_bytes = 0;
    for (auto &k : A.keys) raw_bytes = std::max(raw_bytes, (size_t)a.rows * width_of(a.types[k.second]));
    for (auto &k : B.keys) raw_bytes = std::max(raw_bytes, (size_t)b.rows * width_of(b.types[k.second]));
    uint8_t *raw = nullptr;
    EF_TRY(c.alloc(raw, raw_bytes));
    for (auto *pr : {&a, &b}) {
        const OperandModel &op = pr == &a ? A : B;
        auto &dst = pr == &a ? a_key : b_key;
        for (auto &k : op.keys) {
            EF_TRY(c.alloc(dst[k.first], (size_t)pr->rows));
            EF_TRY(upload_raw(c, *pr, (int)k.second, raw));
            EF_TRY(cast_keys(pr->types[k.second], raw, pr->rows, dst[k.first], stats + k.first, c.stream));
        }
    }
    std::vector<KeyStats> st(nd ? nd : 1);
    EF_TRY(c.download<KeyStats>(st.data(), stats, st.size()));
    EF_TRY(c.sync());
    std::vector<uint64_t> extent(nd, 0);
    for (uint32_t d = 0; d < nd; d++) {
        if (st[d].bad || st[d].lo < 0 || st[d].hi == std::numeric_limits<int64_t>::max()) return EF_DECLINED;
        extent[d] = (uint64_t)(st[d].hi + 1);
    }
    std::vector<uint32_t> wzs = ws;
    wzs.insert(wzs.end(), zs.begin(), zs.end());
    const uint64_t nw = size_of(ws, extent), nx = size_of(xs, extent), nz = size_of(zs, extent),
                   ny = size_of(ys, extent);
    const uint64_t per_batch = nw ? nx / nw : 0;
    const uint64_t out_rows = domain ? (uint64_t)domain->rows : mul_sat(nx, ny);
    if (!worth(mul_sat(nx, nz), a.rows) || !worth(mul_sat(mul_sat(nw, nz), ny), b.rows) ||
        !worth(mul_sat(nx, ny), out_rows))
        return EF_DECLINED;
    const uint64_t na = nx * nz, nb = nw * nz * ny, ncell = nx * ny;
    const uint64_t out_columns = m.outputs.size() + 1 + (domain ? 1 : 0);
    const double working = (double)na * 9 + (double)nx * 2 + (double)nb * 9 + ((double)a.rows + b.rows) * 8 +
                           (double)ncell * 16 + (double)out_rows * out_columns * 8;
    EF_TRY(c.reserve(working, "the dense algorithm's matrices"));

    uint64_t missing_bits;
    std::memcpy(&missing_bits, &m.missing, 8);
    const bool zero_default = m.has_domain && m.has_missing && missing_bits == 0;
    for (auto *pr : {&a, &b}) {
        const OperandModel &op = pr == &a ? A : B;
        if (pr->batches.empty()) continue;
        if (op.factor < 0 || !factor_type_ok(pr->types[op.factor], accept32)) return EF_DECLINED;
    }
    auto dims = [&](const std::vector<uint32_t> &role, const std::vector<int64_t *> &keys) {
        Dims d{};
        d.n = (int)role.size();
        for (size_t i = 0; i < role.size(); i++) {
            d.key[i] = keys[role[i]];
            d.extent[i] = (int64_t)extent[role[i]];
        }
        return d;
    };
    if (xs.size() > MAX_DIMS || wzs.size() > MAX_DIMS || ys.size() > MAX_DIMS || zs.size() > MAX_DIMS ||
        m.outputs.size() > MAX_DIMS || m.conditions.size() > MAX_CONDS) {
        set_error("more dimensions or conditions than the device kernels take");
        return EF_DECLINED;
    }
    // Placement (dense.rs:319-393).
    double *am, *bm, *fa, *fb;
    uint32_t *held, *present;
    uint8_t *reached, *valued = nullptr, *va, *vb;
    PlaceFlags *flags;
    c.begin_zeros();
    EF_TRY(c.zeros(am, na));
    EF_TRY(c.zeros(held, na));
    EF_TRY(c.zeros(reached, nx));
    if (!zero_default) EF_TRY(c.zeros(valued, nx));
    EF_TRY(c.zeros(bm, nb));
    EF_TRY(c.zeros(present, nb));
    EF_TRY(c.zeros(flags, 1));
    EF_TRY(c.end_zeros());
    EF_TRY(c.alloc(fa, (size_t)a.rows));
    EF_TRY(c.alloc(va, (size_t)a.rows));
    EF_TRY(c.alloc(fb, (size_t)b.rows));
    EF_TRY(c.alloc(vb, (size_t)b.rows));
    bool b_nulls = false;
    if (a.rows) EF_TRY(upload_factor(c, a, A.factor, fa, va, nullptr));
    if (b.rows) EF_TRY(upload_factor(c, b, B.factor, fb, vb, &b_nulls));
    EF_TRY(place_a(a.rows, dims(xs, a_key), dims(zs, a_key), nz, fa, va, held, am, reached, valued, flags,
                   c.stream));
    EF_TRY(place_b(b.rows, dims(wzs, b_key), dims(ys, b_key), ny, fb, vb, present, bm, flags, c.stream));
    PlaceFlags pf;
    EF_TRY(c.download<PlaceFlags>(&pf, flags, 1));
    EF_TRY(c.sync());
    if (pf.duplicate) return EF_DECLINED;
    const bool finite = !pf.not_finite;
    // Without duplicates, b is complete when it has one row per cell.
    const bool holes = b_nulls || (uint64_t)b.rows != nb;
    if (holes && (!zero_default || !finite)) return EF_DECLINED;

    std::vector<uint32_t> order = xs;
    order.insert(order.end(), ys.begin(), ys.end());
    Conds conds{};
    conds.n = (int)m.conditions.size();
    for (size_t i = 0; i < m.conditions.size(); i++) {
        const ef_condition &cd = m.conditions[i];
        conds.op[i] = (int)cd.comparison;
        conds.left_in_x[i] = index_in(xs, cd.left) >= 0;
        conds.right_in_x[i] = index_in(xs, cd.right) >= 0;
        conds.left_stride[i] = stride_in(conds.left_in_x[i] ? xs : ys, cd.left, extent);
        conds.right_stride[i] = stride_in(conds.right_in_x[i] ? xs : ys, cd.right, extent);
        conds.left_extent[i] = extent[cd.left] ? extent[cd.left] : 1;
        conds.right_extent[i] = extent[cd.right] ? extent[cd.right] : 1;
        conds.offset[i] = cd.offset;
    }
    // The domain's rows located on cells (dense.rs:400-434, 595-615).
    std::vector<int64_t *> cells;
    if (domain) {
        CellKeys ck{};
        ck.n = (int)m.outputs.size();
        int64_t *dkeys;
        uint8_t *dvalid;
        const size_t nrow = (size_t)domain->rows;
        EF_TRY(c.alloc(dkeys, nrow * m.outputs.size()));
        EF_TRY(c.alloc(dvalid, nrow * m.outputs.size()));
        size_t raw2 = 0;
        for (auto &o : m.outputs) {
            const int col = o.domain_column;
            if (col < 0 || (size_t)col >= domain->columns()) {
                set_error("a domain column is past the domain's columns");
                return EF_INVALID;
            }
            if (!is_integer(domain->types[col]) && !domain->batches.empty()) return EF_DECLINED;
            raw2 = std::max(raw2, nrow * width_of(domain->types[col]));
        }
        uint8_t *rawd;
        EF_TRY(c.alloc(rawd, raw2));
        for (size_t i = 0; i < m.outputs.size(); i++) {
            const ef_output &o = m.outputs[i];
            EF_TRY(upload_validity(c, *domain, o.domain_column, dvalid + i * nrow));
            if (!domain->batches.empty()) {
                EF_TRY(upload_raw(c, *domain, o.domain_column, rawd));
                EF_TRY(widen_keys(domain->types[o.domain_column], (int64_t)nrow, rawd, dvalid + i * nrow,
                                  dkeys + i * nrow, c.stream));
            }
            ck.key[i] = dkeys + i * nrow;
            ck.valid[i] = dvalid + i * nrow;
            ck.extent[i] = (int64_t)extent[o.dimension];
            ck.stride[i] = (int64_t)stride_in(order, o.dimension, extent);
        }
        int64_t *cell_all;
        EF_TRY(c.alloc(cell_all, nrow));
        EF_TRY(locate((int64_t)nrow, ck, cell_all, c.stream));
        if (c.options && c.options->domain_rows == EF_DOMAIN_ROWS_UNIQUE_KEYS) {
            uint32_t *seen;
            int *repeated;
            EF_TRY(c.zeros(seen, ncell));
            EF_TRY(c.zeros(repeated, 1));
            EF_TRY(unique_cells((int64_t)nrow, cell_all, seen, repeated, c.stream));
            int rep = 0;
            EF_TRY(c.download<int>(&rep, repeated, 1));
            EF_TRY(c.sync());
            if (rep) return EF_DECLINED;
        }
        int64_t off = 0;
        for (size_t bi = 0; bi < domain->batches.size(); bi++) {
            cells.push_back(cell_all + off);
            off += domain->batch_rows[bi];
        }
    }

    double *out;
    EF_TRY(c.alloc(out, ncell));
    // Emission (dense.rs:468-577), after the contraction has run: deferred
    // with it when the fold is one of a batch whose MFMA products share a
    // launch.  It holds the call's device buffers by value.
    auto finish_fold = [=, &c, &result, &m, &fields]() -> ef_status {
        if (domain) {
            // dense.rs:488-515: each domain row gets its cell's value where the
            // group exists (and has a value), else the default or NULL.
            for (size_t bi = 0; bi < domain->batches.size(); bi++) {
                const int64_t n = domain->batch_rows[bi];
                double *vals;
                uint8_t *valid;
                EF_TRY(c.alloc(vals, (size_t)n));
                EF_TRY(c.alloc(valid, (size_t)n));
                EF_TRY(domain_values(n, cells[bi], ny, out, zero_default ? nullptr : valued, conds, m.has_missing,
                                     m.missing, vals, valid, c.stream));
                std::vector<OutColumn> cols;
                for (size_t i = 0; i < m.carry.size(); i++)
                    cols.push_back(copy_column(domain->batches[bi][m.carry[i]], n, fields[i].nullable));
                OutColumn v;
                EF_TRY(fetch_column(c, Type::Float64, vals, m.has_missing ? nullptr : valid, n, fields.back().nullable, v));
                cols.push_back(std::move(v));
                result.push_back(std::move(cols));
            }
            return EF_OK;
        }
        // dense.rs:516-577: every cell of every reached row, keys recomputed
        // from positions, filtered by the conditions.
        uint32_t *tmp32, *excl, *rows;
        void *scan_tmp;
        uint8_t *scan_bytes;
        EF_TRY(c.alloc(tmp32, nx));
        EF_TRY(c.alloc(excl, nx));
        EF_TRY(c.alloc(rows, nx));
        EF_TRY(c.alloc(scan_bytes, scan_scratch_bytes(std::max<uint64_t>(nx, ncell))));
        scan_tmp = scan_bytes;
        EF_TRY(reached_rows((int64_t)nx, reached, tmp32, excl, rows, scan_tmp, c.stream));
        uint32_t last[2] = {0, 0};
        uint8_t last_reached = 0;
        if (nx) {
            EF_TRY(c.download<uint32_t>(&last[0], excl + nx - 1, 1));
            EF_TRY(c.download<uint8_t>(&last_reached, reached + nx - 1, 1));
        }
        EF_TRY(c.sync());
        const uint64_t n_rows = nx ? (uint64_t)last[0] + last_reached : 0;
        const uint64_t n = n_rows * ny;
        uint32_t *keep = nullptr, *kexcl = nullptr;
        uint64_t kept = n;
        if (!m.conditions.empty() && n) {
            EF_TRY(c.alloc(keep, n));
            EF_TRY(c.alloc(kexcl, n));
            EF_TRY(keep_cells((int64_t)n, ny, rows, conds, keep, c.stream));
            EF_TRY(scan_u32(keep, kexcl, n, scan_tmp, c.stream));
            uint32_t lk[2];
            EF_TRY(c.download<uint32_t>(&lk[0], kexcl + n - 1, 1));
            EF_TRY(c.download<uint32_t>(&lk[1], keep + n - 1, 1));
            EF_TRY(c.sync());
            kept = (uint64_t)lk[0] + lk[1];
        }
        OutKeys ok{};
        ok.n = (int)m.outputs.size();
        std::vector<uint8_t *> key_out(m.outputs.size());
        for (size_t i = 0; i < m.outputs.size(); i++) {
            const uint32_t d = m.outputs[i].dimension;
            const int w = width_of(fields[i].type);
            EF_TRY(c.alloc(key_out[i], kept * (w ? w : 8)));
            ok.dst[i] = key_out[i];
            ok.width[i] = w;
            ok.in_x[i] = index_in(xs, d) >= 0;
            ok.stride[i] = stride_in(ok.in_x[i] ? xs : ys, d, extent);
            ok.extent[i] = extent[d] ? extent[d] : 1;
        }
        double *vals;
        uint8_t *valid;
        EF_TRY(c.alloc(vals, kept));
        EF_TRY(c.alloc(valid, kept));
        EF_TRY(emit((int64_t)n, ny, rows, keep, kexcl, ok, out, valued, vals, valid, c.stream));
        std::vector<OutColumn> cols;
        for (size_t i = 0; i < m.outputs.size(); i++) {
            OutColumn k;
            EF_TRY(fetch_column(c, fields[i].type, key_out[i], nullptr, (int64_t)kept, fields[i].nullable, k));
            cols.push_back(std::move(k));
        }
        OutColumn v;
        EF_TRY(fetch_column(c, Type::Float64, vals, valid, (int64_t)kept, true, v));
        cols.push_back(std::move(v));
        result.push_back(std::move(cols));
        return EF_OK;
    };
    // The contraction (dense.rs:706-770): in the reference's order, or on
    // MFMA where the caller asked for it and every factor is finite (the
    // matrices hold 0 for absent and NULL cells, which then add nothing;
    // with an infinite factor, 0 * inf would).
    const uint64_t pb = per_batch ? per_batch : 1;
    const bool all_finite = finite && !pf.b_not_finite;
    const uint32_t prec = c.options ? c.options->precision : EF_PRECISION_FP64;
    const bool f32 = a.rows && b.rows && a.types[A.factor] == Type::Float32 && b.types[B.factor] == Type::Float32;
    if (ncell == 0) {
    } else if (f32 && prec != EF_PRECISION_FP64 && all_finite) {
        // Float32 factors, FP32-accumulated products on the matrix cores:
        // one grouped launch of the W batches' products (per_batch x nz)
        // (nz x ny), its tile shape from the autotuner keyed on (M, N, K,
        // key density); the placed Float64 values narrow back exactly.
        float *a32, *b32, *c32;
        EF_TRY(c.alloc(a32, na));
        EF_TRY(c.alloc(b32, nb));
        EF_TRY(c.alloc(c32, ncell));
        EF_TRY(narrow_f32(am, (int64_t)na, a32, c.stream));
        EF_TRY(narrow_f32(bm, (int64_t)nb, b32, c.stream));
        const double density = std::min((double)a.rows / (double)std::max<uint64_t>(na, 1),
                                        (double)b.rows / (double)std::max<uint64_t>(nb, 1));
        std::vector<ef_gemm_problem> probs(nw);
        for (uint64_t w = 0; w < nw; w++)
            probs[w] = {a32 + w * pb * nz, b32 + w * nz * ny, c32 + w * pb * ny, (size_t)pb, (size_t)ny, (size_t)nz};
        if (batch) {
            // The products join the batch's one launch; the widening and
            // the emission wait for it.
            batch->problems.insert(batch->problems.end(), probs.begin(), probs.end());
            batch->density.insert(batch->density.end(), probs.size(), density);
            batch->precision = prec;
            batch->finish = [=, &c]() -> ef_status {
                EF_TRY(to_f64(Type::Float32, c32, (int64_t)ncell, out, c.stream));
                return finish_fold();
            };
            return EF_OK;
        }
        std::vector<double> dens(nw, density);
        EF_TRY(grouped_run_once(probs.data(), probs.size(), prec, dens.data(), c.stream));
        EF_TRY(to_f64(Type::Float32, c32, (int64_t)ncell, out, c.stream));
    } else if (!f32 && c.options && c.options->fp64_mfma && all_finite && nw <= 65535) {
        // Float64 factors on FP64 MFMA (opted into): another summation
        // order, within 2 (n - 1) u sum|terms| of the exact one.
        EF_TRY(launch_dense_product_f64(am, bm, out, nw, pb, ny, nz, c.stream));
    } else {
        // The predicated kernel skips the cells `held` does not mark valued
        // with a select.  When every b is finite, those cells' zero factors
        // give +-0 products instead, and adding +-0 leaves a sum from +0.0 as
        // it is (such a sum is never -0.0): the same bits, without the flags.
        EF_TRY(launch_dense_accumulate_f64(am, held, bm, nx, nz, ny, pb, pf.b_not_finite ? 1 : 0, out, c.stream));
    }

    return finish_fold();
}

// ---------------------------------------------------------------------
// The discrimination algorithm (discriminate.rs:240-650)
// ---------------------------------------------------------------------

namespace {

// A key column on the device: the rows of one or more relations' columns
// concatenated, as raw values and a validity byte per row.
struct DevColumn {
    Type type = Type::Other;
    uint8_t *raw = nullptr;
    uint8_t *valid = nullptr;
    int64_t n = 0;
    bool nulls = false;
};

ef_status upload_column(Call &c, const std::vector<std::pair<const Relation *, int>> &parts, bool zero_signs_equal,
                        DevColumn &out) {
    out = DevColumn();
    out.type = parts[0].first->types[parts[0].second];
    for (auto &p : parts) {
        if (p.first->types[p.second] != out.type) {
            set_error("a dimension's key columns have two types");
            return EF_INVALID;
        }
        out.n += p.first->rows;
        for (auto &bt : p.first->batches)
            if (bt[p.second].null_count) out.nulls = true;
    }
    const int w = width_of(out.type);
    if (!w) {
        set_error("keys must be integers or floats");
        return EF_INVALID;
    }
    EF_TRY(c.alloc(out.raw, (size_t)out.n * w));
    EF_TRY(c.alloc(out.valid, (size_t)out.n));
    int64_t off = 0;
    for (auto &p : parts) {
        EF_TRY(upload_raw(c, *p.first, p.second, out.raw + off * w));
        EF_TRY(upload_validity(c, *p.first, p.second, out.valid + off));
        off += p.first->rows;
    }
    if (zero_signs_equal && is_float(out.type)) EF_TRY(zero_signs(out.type, out.raw, out.n, c.stream));
    return EF_OK;
}

// rank_columns (discriminate.rs:240-277): each row's dense rank in the
// key order of `cols` (first most significant, NULL first), and each
// rank's lowest row.  A column is ranked by its order words; a nullable
// one again by 0 for NULL, 2r + 1 otherwise; a later column joins the
// running rank as (rank << 32 | its rank), ranked again.
ef_status rank_columns(Call &c, const std::vector<DevColumn *> &cols, int64_t n, uint32_t *rank, uint32_t *first,
                       size_t &k, void *scratch) {
    k = 0;
    if (n == 0) return EF_OK;
    if (cols.empty()) {
        EF_HIP(hipMemsetAsync(rank, 0, (size_t)n * 4, c.stream));
        EF_HIP(hipMemsetAsync(first, 0, 4, c.stream));
        k = 1;
        return EF_OK;
    }
    uint64_t *w;
    uint32_t *r;
    EF_TRY(c.alloc(w, (size_t)n));
    EF_TRY(c.alloc(r, (size_t)n));
    for (size_t ci = 0; ci < cols.size(); ci++) {
        const DevColumn &col = *cols[ci];
        uint32_t *dst = ci == 0 ? rank : r;
        EF_TRY(order_words(col.type, col.raw, col.nulls ? col.valid : nullptr, n, w, c.stream));
        EF_TRY(rank_u64(w, (size_t)n, dst, first, &k, scratch, c.stream));
        if (col.nulls) {
            EF_TRY(null_codes(dst, col.valid, n, w, c.stream));
            EF_TRY(rank_u64(w, (size_t)n, dst, first, &k, scratch, c.stream));
        }
        if (ci > 0) {
            EF_TRY(combine_ranks(rank, r, n, w, c.stream));
            EF_TRY(rank_u64(w, (size_t)n, rank, first, &k, scratch, c.stream));
        }
    }
    return EF_OK;
}

// An operand's joining rows in (hi, lo, factor) order (discriminate.rs:
// 493-506): stable sorts by the pair of ranks, and where a pair repeats,
// first by the factor's bits and NULL first (break_ties), all stable from
// row order.  Returns the sorted words (hi << 32 | lo) and rows.
ef_status sort_rows(Call &c, const uint32_t *rows, int64_t n, const uint32_t *hi, const uint32_t *lo,
                    const uint8_t *valid, const double *value, uint64_t *w_sorted, uint32_t *r_sorted,
                    void *scratch) {
    if (n == 0) return EF_OK;
    uint64_t *w;
    EF_TRY(c.alloc(w, (size_t)n));
    EF_TRY(pair_words(rows, n, hi, lo, w, c.stream));
    EF_TRY(sort_u64(w, rows, (size_t)n, w_sorted, r_sorted, scratch, c.stream));
    int *flag;
    EF_TRY(c.alloc(flag, 1));
    bool repeats = false;
    EF_TRY(any_repeated(w_sorted, n, flag, &repeats, c.stream));
    if (!repeats) return EF_OK;
    uint64_t *bits, *vw;
    uint32_t *t1, *t2;
    EF_TRY(c.alloc(bits, (size_t)n));
    EF_TRY(c.alloc(vw, (size_t)n));
    EF_TRY(c.alloc(t1, (size_t)n));
    EF_TRY(c.alloc(t2, (size_t)n));
    EF_TRY(tie_words(rows, n, valid, value, bits, vw, c.stream));
    EF_TRY(sort_u64(bits, rows, (size_t)n, nullptr, t1, scratch, c.stream));
    EF_TRY(tie_words(t1, n, valid, value, bits, vw, c.stream));
    EF_TRY(sort_u64(vw, t1, (size_t)n, nullptr, t2, scratch, c.stream));
    EF_TRY(pair_words(t2, n, hi, lo, w, c.stream));
    return sort_u64(w, t2, (size_t)n, w_sorted, r_sorted, scratch, c.stream);
}

uint64_t column_bytes(const Relation &r, int col) {
    uint64_t b = (uint64_t)r.rows * width_of(r.types[col]);
    for (auto &bt : r.batches)
        if (bt[col].validity) b += (uint64_t)(bt[col].length + 7) / 8;
    return b;
}

uint64_t relation_bytes(const Relation &r) {
    uint64_t b = 0;
    for (size_t col = 0; col < r.columns(); col++) b += column_bytes(r, (int)col);
    return b;
}

}  // namespace

ef_status fold_discriminate_impl(const FoldModel &m, const Relation &a, const Relation &b, const Relation *domain,
                                 const std::vector<OutField> &fields, Call &c,
                                 std::vector<std::vector<OutColumn>> &result) {
    const bool accept32 = c.options && c.options->accept_float32;
    const bool zeros_equal = c.options && c.options->float_keys == EF_FLOAT_KEYS_SIGNED_ZEROS_EQUAL;
    const int64_t na = a.rows, nb = b.rows, nd = domain ? domain->rows : 0;
    const uint64_t ranked_rows = std::max<uint64_t>(na + nb, std::max(na, nb) + nd);
    if (ranked_rows > U32_MAX) {
        set_error("EinFold's discrimination algorithm ranks at most 2^32 - 1 rows together");
        return EF_INVALID;
    }
    if (m.operands.size() != 2) {
        set_error("the discrimination algorithm takes two operands");
        return EF_INVALID;
    }
    const OperandModel &A = m.operands[0], &B = m.operands[1];
    std::vector<uint32_t> shared;
    for (uint32_t d = 0; d < m.equality.size(); d++)
        if (A.column_of(d) >= 0 && B.column_of(d) >= 0) shared.push_back(d);
    std::vector<size_t> a_out, b_out;
    for (size_t i = 0; i < m.outputs.size(); i++)
        (A.column_of(m.outputs[i].dimension) >= 0 ? a_out : b_out).push_back(i);
    if (A.factor < 0 || (size_t)A.factor >= a.columns() || !factor_type_ok(a.types[A.factor], accept32) ||
        B.factor < 0 || (size_t)B.factor >= b.columns() || !factor_type_ok(b.types[B.factor], accept32)) {
        set_error("factors must be Float64");
        return EF_INVALID;
    }
    uint64_t key_bytes = 0;
    for (auto &k : A.keys) key_bytes += column_bytes(a, (int)k.second);
    for (auto &k : B.keys) key_bytes += column_bytes(b, (int)k.second);
    const double ranked = ((double)na + nb) * 2 + (double)nd * 2;
    EF_TRY(c.reserve(ranked * 28 + ((double)na + nb) * 40 + (double)key_bytes * 2,
                     "the discrimination algorithm's ranks"));

    void *rscratch;
    uint8_t *rs_bytes;
    EF_TRY(c.alloc(rs_bytes, rank_scratch_bytes(std::max<uint64_t>(ranked_rows, 1))));
    rscratch = rs_bytes;
    double *fa, *fb;
    uint8_t *va, *vb;
    EF_TRY(c.alloc(fa, (size_t)na));
    EF_TRY(c.alloc(va, (size_t)na));
    EF_TRY(c.alloc(fb, (size_t)nb));
    EF_TRY(c.alloc(vb, (size_t)nb));
    if (na) EF_TRY(upload_factor(c, a, A.factor, fa, va, nullptr));
    if (nb) EF_TRY(upload_factor(c, b, B.factor, fb, vb, nullptr));

    // The join's keys over both operands; rows with a NULL in a key
    // compared with `=` join nothing.
    std::deque<DevColumn> zcols;
    std::vector<DevColumn *> zptr;
    uint8_t *a_join, *b_join;
    EF_TRY(c.alloc(a_join, (size_t)na));
    EF_TRY(c.alloc(b_join, (size_t)nb));
    if (na) EF_HIP(hipMemsetAsync(a_join, 1, (size_t)na, c.stream));
    if (nb) EF_HIP(hipMemsetAsync(b_join, 1, (size_t)nb, c.stream));
    for (uint32_t d : shared) {
        zcols.emplace_back();
        EF_TRY(upload_column(c, {{&a, A.column_of(d)}, {&b, B.column_of(d)}}, zeros_equal, zcols.back()));
        zptr.push_back(&zcols.back());
        if (m.equality[d] == EF_EQUAL && zcols.back().nulls) {
            EF_TRY(and_valid(a_join, zcols.back().valid, na, c.stream));
            EF_TRY(and_valid(b_join, zcols.back().valid + na, nb, c.stream));
        }
    }
    uint32_t *z_rank, *z_first;
    EF_TRY(c.alloc(z_rank, (size_t)(na + nb)));
    EF_TRY(c.alloc(z_first, (size_t)(na + nb)));
    size_t kz = 0;
    EF_TRY(rank_columns(c, zptr, na + nb, z_rank, z_first, kz, rscratch));

    // Output key columns to rank: an operand's own values, followed by the
    // domain's values for the same outputs when a domain is given.
    std::deque<DevColumn> xcols, ycols;
    std::vector<DevColumn *> xptr, yptr;
    for (int side = 0; side < 2; side++) {
        const auto &outs = side == 0 ? a_out : b_out;
        const Relation &r = side == 0 ? a : b;
        const OperandModel &op = side == 0 ? A : B;
        auto &cols = side == 0 ? xcols : ycols;
        auto &ptrs = side == 0 ? xptr : yptr;
        for (size_t i : outs) {
            std::vector<std::pair<const Relation *, int>> parts = {{&r, op.column_of(m.outputs[i].dimension)}};
            if (nd) {
                const int dc = m.outputs[i].domain_column;
                if (dc < 0 || (size_t)dc >= domain->columns()) {
                    set_error("a domain column is past the domain's columns");
                    return EF_INVALID;
                }
                parts.push_back({domain, dc});
            }
            cols.emplace_back();
            EF_TRY(upload_column(c, parts, zeros_equal, cols.back()));
            ptrs.push_back(&cols.back());
        }
    }
    uint32_t *x_rank, *x_first, *y_rank, *y_first;
    EF_TRY(c.alloc(x_rank, (size_t)(na + nd)));
    EF_TRY(c.alloc(x_first, (size_t)(na + nd)));
    EF_TRY(c.alloc(y_rank, (size_t)(nb + nd)));
    EF_TRY(c.alloc(y_first, (size_t)(nb + nd)));
    size_t kx = 0, ky = 0;
    EF_TRY(rank_columns(c, xptr, na + nd, x_rank, x_first, kx, rscratch));
    EF_TRY(rank_columns(c, yptr, nb + nd, y_rank, y_first, ky, rscratch));
    const int64_t nx = (int64_t)std::max<size_t>(kx, 1), nz = (int64_t)std::max<size_t>(kz, 1),
                  ny = (int64_t)std::max<size_t>(ky, 1);

    // Sort the rows that join: a by (x rank, z rank), b by (z rank, y rank);
    // rows whose ranks tie are ordered by their factor bits.
    uint32_t *flags32, *excl32, *rows_a, *rows_b;
    uint8_t *scan_bytes;
    const int64_t nmax = std::max<int64_t>(std::max(na, nb), 1);
    EF_TRY(c.alloc(flags32, (size_t)nmax));
    EF_TRY(c.alloc(excl32, (size_t)nmax));
    EF_TRY(c.alloc(rows_a, (size_t)na));
    EF_TRY(c.alloc(rows_b, (size_t)nb));
    EF_TRY(c.alloc(scan_bytes, scan_scratch_bytes((size_t)nmax)));
    uint32_t count_a = 0, count_b = 0;
    EF_TRY(select_rows(a_join, na, flags32, excl32, rows_a, scan_bytes, &count_a, c.stream));
    EF_TRY(select_rows(b_join, nb, flags32, excl32, rows_b, scan_bytes, &count_b, c.stream));
    uint64_t *wa, *wb;
    uint32_t *ra, *rb;
    EF_TRY(c.alloc(wa, count_a));
    EF_TRY(c.alloc(ra, count_a));
    EF_TRY(c.alloc(wb, count_b));
    EF_TRY(c.alloc(rb, count_b));
    EF_TRY(sort_rows(c, rows_a, count_a, x_rank, z_rank, va, fa, wa, ra, rscratch));
    EF_TRY(sort_rows(c, rows_b, count_b, z_rank + na, y_rank, vb, fb, wb, rb, rscratch));
    uint32_t *a_start, *b_start, *a_z, *b_y;
    uint8_t *a_valid, *b_valid;
    double *a_value, *b_value;
    EF_TRY(c.alloc(a_start, (size_t)nx + 1));
    EF_TRY(c.alloc(b_start, (size_t)nz + 1));
    EF_TRY(c.alloc(a_z, count_a));
    EF_TRY(c.alloc(a_valid, count_a));
    EF_TRY(c.alloc(a_value, count_a));
    EF_TRY(c.alloc(b_y, count_b));
    EF_TRY(c.alloc(b_valid, count_b));
    EF_TRY(c.alloc(b_value, count_b));
    EF_TRY(segment_starts(wa, count_a, nx, a_start, c.stream));
    EF_TRY(segment_starts(wb, count_b, nz, b_start, c.stream));
    EF_TRY(gather_sorted(ra, wa, count_a, va, fa, a_z, a_valid, a_value, c.stream));
    EF_TRY(gather_sorted(rb, wb, count_b, vb, fb, b_y, b_valid, b_value, c.stream));

    // The groups: no more than the joined pairs, nor than every (x, y)
    // (discriminate.rs:530-545), and per x no more than min(ny, its pairs).
    uint64_t *bound, *bound_off;
    unsigned long long *pairs_d;
    EF_TRY(c.alloc(bound, (size_t)nx));
    EF_TRY(c.alloc(bound_off, (size_t)nx));
    EF_TRY(c.zeros(pairs_d, 1));
    EF_TRY(group_bounds(a_start, a_z, b_start, nx, ny, bound, pairs_d, c.stream));
    uint8_t *scan64;
    EF_TRY(c.alloc(scan64, scan64_scratch_bytes((size_t)nx)));
    EF_TRY(scan_u64(bound, bound_off, (size_t)nx, scan64, c.stream));
    unsigned long long pairs = 0;
    uint64_t last_off = 0, last_bound = 0;
    EF_TRY(c.download<unsigned long long>(&pairs, pairs_d, 1));
    EF_TRY(c.download<uint64_t>(&last_off, bound_off + nx - 1, 1));
    EF_TRY(c.download<uint64_t>(&last_bound, bound + nx - 1, 1));
    EF_TRY(c.sync());
    const double groups = std::min<double>((double)pairs, (double)nx * (double)ny);
    const double output = m.has_domain ? (double)nd * 8 : groups * 8 * (double)(m.outputs.size() + 1);
    EF_TRY(c.reserve(groups * 24 + output, "the discrimination algorithm's groups"));
    const uint64_t staged = last_off + last_bound;
    uint32_t *g_y, *g_count;
    double *g_sum;
    uint8_t *g_valued;
    EF_TRY(c.alloc(g_y, staged));
    EF_TRY(c.alloc(g_sum, staged));
    EF_TRY(c.alloc(g_valued, staged));
    EF_TRY(c.alloc(g_count, (size_t)nx));

    size_t free_b = 0, total_b = 0;
    EF_HIP(hipMemGetInfo(&free_b, &total_b));
    // Heavy join ranks: where every (x, y) cell fits in HBM and the groups
    // fill at least a quarter of them, the z's whose b bucket holds
    // HEAVY_ROWS rows or more (the HEAVY_MAX largest) run as dense phases
    // (gustavson_heavy): the same sums in the same order.
    constexpr uint32_t HEAVY_ROWS = 1024;
    constexpr size_t HEAVY_MAX = 16;
    std::vector<uint32_t> heavy, heavy_rows;
    const double cells = (double)nx * (double)ny;
    if (ny > 512 && nz > 0 && cells * 9 <= std::min((double)free_b / 2, 8.0 * (double)(1ull << 30)) &&
        groups >= cells / 4) {
        std::vector<uint32_t> bst((size_t)nz + 1);
        EF_TRY(c.download<uint32_t>(bst.data(), b_start, (size_t)nz + 1));
        EF_TRY(c.sync());
        std::vector<std::pair<uint32_t, uint32_t>> big;   // (rows, z)
        for (uint64_t z = 0; z < (uint64_t)nz; z++)
            if (bst[z + 1] - bst[z] >= HEAVY_ROWS) big.push_back({bst[z + 1] - bst[z], (uint32_t)z});
        std::sort(big.begin(), big.end(), [](const auto &l, const auto &r) {
            return l.first != r.first ? l.first > r.first : l.second < r.second;
        });
        if (big.size() > HEAVY_MAX) big.resize(HEAVY_MAX);
        std::sort(big.begin(), big.end(), [](const auto &l, const auto &r) { return l.second < r.second; });
        for (auto &h : big) {
            heavy.push_back(h.second);
            heavy_rows.push_back(h.first);
        }
    }

    // Gustavson's product: wave slots, each with an accumulator of ny.
    const double per_slot = (double)ny * 9.0;
    int64_t slots = std::min<int64_t>(2048, (int64_t)std::max(4.0, std::floor((double)free_b / 4 / per_slot)));
    slots = std::min<int64_t>(slots, ((nx + 3) / 4) * 4);
    slots = std::max<int64_t>(4, (slots / 4) * 4);
    GustavsonArgs g{};
    g.a_start = a_start;
    g.a_z = a_z;
    g.a_valid = a_valid;
    g.a_value = a_value;
    g.b_start = b_start;
    g.b_y = b_y;
    g.b_valid = b_valid;
    g.b_value = b_value;
    g.nx = nx;
    g.ny = ny;
    g.bound_off = bound_off;
    g.g_y = g_y;
    g.g_sum = g_sum;
    g.g_valued = g_valued;
    g.g_count = g_count;
    if (!heavy.empty()) {
        c.begin_zeros();
        EF_TRY(c.zeros(g.acc, (size_t)(nx * ny)));
        EF_TRY(c.zeros(g.state, (size_t)(nx * ny)));
        EF_TRY(c.end_zeros());
        std::vector<uint64_t> run_off(heavy.size());
        uint64_t run_total = 0;
        for (size_t k = 0; k < heavy.size(); k++) {
            run_off[k] = run_total;
            run_total += heavy_rows[k];
        }
        uint32_t *heavy_dev, *split, *runs, *n_runs;
        uint64_t *run_off_dev;
        EF_TRY(c.alloc(heavy_dev, heavy.size()));
        EF_TRY(c.upload<uint32_t>(heavy_dev, heavy.data(), heavy.size()));
        EF_TRY(c.alloc(run_off_dev, heavy.size()));
        EF_TRY(c.upload<uint64_t>(run_off_dev, run_off.data(), heavy.size()));
        EF_TRY(c.alloc(split, 2 * heavy.size() * (size_t)nx));
        EF_TRY(c.alloc(runs, (size_t)run_total));
        EF_TRY(c.alloc(n_runs, heavy.size()));
        EF_TRY(gustavson_heavy(g, heavy.data(), heavy_dev, heavy_rows.data(), (int)heavy.size(), split, runs,
                               run_off.data(), run_off_dev, n_runs, c.stream));
    } else {
        EF_TRY(c.zeros(g.acc, (size_t)(slots * ny)));
        EF_TRY(c.zeros(g.state, (size_t)(slots * ny)));
        EF_TRY(c.zeros(g.next, 1));
        EF_TRY(gustavson(g, (int)(slots / 4), c.stream));
    }

    if (!m.has_domain) {
        uint64_t *counts64, *out_off;
        EF_TRY(c.alloc(counts64, (size_t)nx));
        EF_TRY(c.alloc(out_off, (size_t)nx));
        EF_TRY(widen_counts(g_count, nx, counts64, c.stream));
        EF_TRY(scan_u64(counts64, out_off, (size_t)nx, scan64, c.stream));
        uint64_t lo = 0, lc = 0;
        EF_TRY(c.download<uint64_t>(&lo, out_off + nx - 1, 1));
        EF_TRY(c.download<uint64_t>(&lc, counts64 + nx - 1, 1));
        EF_TRY(c.sync());
        const uint64_t total = lo + lc;
        GroupKeys keys{};
        keys.n = (int)m.outputs.size();
        if (keys.n > 8) {
            set_error("more than 8 outputs");
            return EF_INVALID;
        }
        std::vector<uint8_t *> kd(keys.n), kv(keys.n);
        for (size_t i = 0; i < m.outputs.size(); i++) {
            auto ia = std::find(a_out.begin(), a_out.end(), i);
            const DevColumn *src = ia != a_out.end() ? xptr[ia - a_out.begin()]
                                                     : yptr[std::find(b_out.begin(), b_out.end(), i) - b_out.begin()];
            const int w = width_of(src->type);
            EF_TRY(c.alloc(kd[i], total * w));
            EF_TRY(c.alloc(kv[i], total));
            keys.src[i] = src->raw;
            keys.src_valid[i] = src->valid;
            keys.dst[i] = kd[i];
            keys.dst_valid[i] = kv[i];
            keys.width[i] = w;
            keys.from_a[i] = ia != a_out.end();
        }
        double *vals;
        uint8_t *valid;
        EF_TRY(c.alloc(vals, total));
        EF_TRY(c.alloc(valid, total));
        uint8_t *emit_scratch;
        EF_TRY(c.alloc(emit_scratch, emit_scratch_bytes((int64_t)ky)));
        EF_TRY(emit_groups(nx, (int64_t)ky, bound_off, out_off, g_count, g_y, g_sum, g_valued, x_first, y_first, keys,
                           emit_scratch, vals, valid, c.stream));
        std::vector<OutColumn> cols;
        for (size_t i = 0; i < m.outputs.size(); i++) {
            OutColumn k;
            EF_TRY(fetch_column(c, fields[i].type, kd[i], kv[i], (int64_t)total, fields[i].nullable, k));
            cols.push_back(std::move(k));
        }
        OutColumn v;
        EF_TRY(fetch_column(c, Type::Float64, vals, valid, (int64_t)total, true, v));
        cols.push_back(std::move(v));
        result.push_back(std::move(cols));
        return EF_OK;
    }
    if (!domain) {
        set_error("a fold with a domain was given no domain relation");
        return EF_INVALID;
    }
    // Domain rows find their group by rank (discriminate.rs:611-648).
    uint8_t *matches;
    EF_TRY(c.alloc(matches, (size_t)nd));
    if (nd) EF_HIP(hipMemsetAsync(matches, 1, (size_t)nd, c.stream));
    for (size_t j = 0; j < a_out.size(); j++)
        if (m.outputs[a_out[j]].domain_equality == EF_EQUAL) EF_TRY(and_valid(matches, xptr[j]->valid + na, nd, c.stream));
    for (size_t j = 0; j < b_out.size(); j++)
        if (m.outputs[b_out[j]].domain_equality == EF_EQUAL) EF_TRY(and_valid(matches, yptr[j]->valid + nb, nd, c.stream));
    int64_t off = 0;
    for (size_t bi = 0; bi < domain->batches.size(); bi++) {
        const int64_t n = domain->batch_rows[bi];
        double *vals;
        uint8_t *valid;
        EF_TRY(c.alloc(vals, (size_t)n));
        EF_TRY(c.alloc(valid, (size_t)n));
        EF_TRY(domain_lookup(n, matches + off, x_rank + na + off, y_rank + nb + off, nx, bound_off, g_count, g_y, g_sum,
                             g_valued, m.has_missing, m.missing, vals, valid, c.stream));
        std::vector<OutColumn> cols;
        for (size_t i = 0; i < m.carry.size(); i++)
            cols.push_back(copy_column(domain->batches[bi][m.carry[i]], n, fields[i].nullable));
        OutColumn v;
        EF_TRY(fetch_column(c, Type::Float64, vals, m.has_missing ? nullptr : valid, n, fields.back().nullable, v));
        cols.push_back(std::move(v));
        result.push_back(std::move(cols));
        off += n;
    }
    return EF_OK;
}

// ---------------------------------------------------------------------
// The fused pipeline (fused.rs:151-489)
// ---------------------------------------------------------------------

namespace {

// Each operand's keys as i64 per dimension (every batch in one array),
// with the dimensions' extents; false if dense keys don't hold.
struct KeyedOperand {
    const Relation *rel = nullptr;
    const OperandModel *op = nullptr;
    std::vector<int64_t *> key;        // by dimension
};

ef_status cast_operand_keys(Call &c, KeyedOperand &ko, size_t n_dims, KeyStats *stats) {
    ko.key.assign(n_dims, nullptr);
    const Relation &r = *ko.rel;
    size_t raw_bytes = 0;
    for (auto &k : ko.op->keys) {
        if (!r.batches.empty() && !is_integer(r.types[k.second])) return EF_DECLINED;
        for (auto &bt : r.batches)
            if (bt[k.second].null_count) return EF_DECLINED;
        raw_bytes = std::max(raw_bytes, (size_t)r.rows * width_of(r.types[k.second]));
    }
    uint8_t *raw;
    EF_TRY(c.alloc(raw, raw_bytes));
    for (auto &k : ko.op->keys) {
        EF_TRY(c.alloc(ko.key[k.first], (size_t)r.rows));
        if (!r.rows) continue;
        EF_TRY(upload_raw(c, r, (int)k.second, raw));
        EF_TRY(cast_keys(r.types[k.second], raw, r.rows, ko.key[k.first], stats + k.first, c.stream));
    }
    return EF_OK;
}

// Whether a postfix program is well formed for the device evaluator: every
// operator finds its operands, the stack never holds more than 16 values,
// and one value is left.
// Whether a score program is a squared difference of one q-side and one
// k-side column, (c0 - c1)^2, followed by up to four operations with a
// literal (negation; +, -, * or / by a literal on either side), and its
// parts in `pa`: the kernels then evaluate it without the interpreter, the
// same operations in the same order.
bool squared_difference(const std::vector<ef_expr_op> &e, PipeArgs &pa) {
    if (e.size() < 5 || e[0].code != EF_EXPR_COLUMN || e[1].code != EF_EXPR_COLUMN || e[2].code != EF_EXPR_SUB ||
        e[3].code != EF_EXPR_LITERAL || e[3].value != 2.0 || e[4].code != EF_EXPR_POW)
        return false;
    if (e[0].operand == e[1].operand || e[0].operand > 1 || e[1].operand > 1) return false;
    int n = 0;
    for (size_t i = 5; i < e.size();) {
        if (n == 4) return false;
        if (e[i].code == EF_EXPR_NEG) {
            pa.post_code[n] = EF_EXPR_NEG;
            pa.post_value[n] = 0.0;
            pa.post_literal_first[n] = 0;
            n++;
            i++;
            continue;
        }
        if (i + 1 < e.size() && e[i].code == EF_EXPR_LITERAL &&
            (e[i + 1].code == EF_EXPR_MUL || e[i + 1].code == EF_EXPR_DIV || e[i + 1].code == EF_EXPR_ADD ||
             e[i + 1].code == EF_EXPR_SUB)) {
            pa.post_code[n] = e[i + 1].code;
            pa.post_value[n] = e[i].value;
            pa.post_literal_first[n] = 0;
            n++;
            i += 2;
            continue;
        }
        return false;
    }
    pa.n_post = n;
    // Monotone in the square: finite literals, no product or quotient by 0
    // (and a finite, nonzero weight scale, checked by the caller).
    pa.sq_mono = 1;
    for (int i = 0; i < n; i++) {
        const double c = pa.post_value[i];
        if (!std::isfinite(c) || ((pa.post_code[i] == EF_EXPR_MUL || pa.post_code[i] == EF_EXPR_DIV) && c == 0.0))
            pa.sq_mono = 0;
    }
    pa.sq_neg_div = n == 2 && pa.post_code[0] == EF_EXPR_NEG && pa.post_code[1] == EF_EXPR_DIV;
    pa.sq_div_pow2 = 0;
    pa.sq_neg_recip = 0.0;
    if (pa.sq_neg_div && std::isfinite(pa.post_value[1]) && pa.post_value[1] != 0.0) {
        int ex = 0;
        const double c = -pa.post_value[1], r = 1.0 / c;
        if (std::frexp(std::fabs(c), &ex) == 0.5 && std::isnormal(r) && r * c == 1.0) {
            pa.sq_div_pow2 = 1;
            pa.sq_neg_recip = r;
        }
    }
    pa.sq_swap = e[0].operand == 1 ? 1 : 0;
    pa.sq_q = (int)(e[0].operand == 0 ? e[0].column : e[1].column);
    pa.sq_k = (int)(e[0].operand == 1 ? e[0].column : e[1].column);
    return true;
}

bool program_ok(const ef_expr_op *ops, size_t n) {
    int depth = 0;
    for (size_t i = 0; i < n; i++) {
        const uint32_t code = ops[i].code;
        if (code == EF_EXPR_COLUMN || code == EF_EXPR_LITERAL) depth++;
        else if (code >= EF_EXPR_NEG && code <= EF_EXPR_LN) { if (depth < 1) return false; }
        else if (code >= EF_EXPR_ADD && code <= EF_EXPR_POW) { if (depth < 2) return false; depth--; }
        else return false;
        if (depth > 16) return false;
    }
    return depth == 1;
}

struct Placement {
    int operand;          // 0 q, 1 k, 2 v
    int column;
    int slot;             // index among the operand's placed columns
};

}  // namespace

ef_status pipeline_impl(const ef_pipeline_desc *pd, const FoldModel &prod, const Relation &q, const Relation &k,
                        const Relation &v, const std::vector<OutField> &fields, Call &c,
                        std::vector<std::vector<OutColumn>> &result, uint64_t &tiles_visited, uint64_t &tiles) {
    const bool accept32 = c.options && c.options->accept_float32;
    if (prod.operands.size() != 2) return EF_DECLINED;
    const OperandModel &oq = prod.operands[0], &ok_ = prod.operands[1];
    OperandModel ov;
    ov.factor = pd->values.factor;
    for (uint32_t i = 0; i < pd->values.n_keys; i++)
        ov.keys.emplace_back(pd->values.keys[i].dimension, pd->values.keys[i].column);
    const uint32_t own = (uint32_t)prod.equality.size();
    const uint32_t n = own + pd->n_dimensions;
    std::vector<uint32_t> outs(pd->outputs, pd->outputs + pd->n_outputs);
    auto output = [&](uint32_t d) { return std::find(outs.begin(), outs.end(), d) != outs.end(); };
    auto holds = [](const OperandModel &o, uint32_t d) { return o.column_of(d) >= 0; };
    // Roles (fused.rs:171-189): W batches, X queries, E features, J keys.
    std::vector<uint32_t> ws, xs, es, js, ys;
    for (uint32_t d = 0; d < own; d++) {
        const bool produced = prod.output_of(d) >= 0;
        const int key = (holds(oq, d) << 4) | (holds(ok_, d) << 3) | (produced << 2) | (output(d) << 1) | holds(ov, d);
        if (key == 0b11111) ws.push_back(d);
        else if (key == 0b10110) xs.push_back(d);
        else if (key == 0b11000) es.push_back(d);
        else if (key == 0b01101) js.push_back(d);
        else return EF_DECLINED;
    }
    for (uint32_t d = own; d < n; d++) {
        if (!output(d)) return EF_DECLINED;
        ys.push_back(d);
    }
    // Masks: conditions between the one X and the one J dimension
    // (fused.rs:190-215), mirrored to read x <cmp> j + offset.
    std::vector<ef_condition> mask;
    for (const ef_condition &cd : prod.conditions) {
        if (xs.size() == 1 && js.size() == 1 && cd.left == xs[0] && cd.right == js[0]) {
            mask.push_back(cd);
        } else if (xs.size() == 1 && js.size() == 1 && cd.left == js[0] && cd.right == xs[0]) {
            if (cd.offset == std::numeric_limits<int64_t>::min()) return EF_DECLINED;
            ef_condition mc = cd;
            mc.left = cd.right;
            mc.right = cd.left;
            mc.offset = -cd.offset;
            static const uint32_t mirror[4] = {EF_GREATER, EF_GREATER_OR_EQUAL, EF_LESS, EF_LESS_OR_EQUAL};
            mc.comparison = mirror[cd.comparison];
            mask.push_back(mc);
        } else {
            return EF_DECLINED;
        }
    }
    const std::vector<uint32_t> local_xs = xs, local_js = js;
    xs.insert(xs.begin(), ws.begin(), ws.end());
    js.insert(js.begin(), ws.begin(), ws.end());
    // What the producer reads (fused.rs:229-262).
    std::vector<Placement> reads;
    std::vector<ef_expr_op> score(prod.row);
    int nq_cols = 0, nk_cols = 0;
    if (prod.row.empty()) {
        if (oq.factor < 0 || ok_.factor < 0) return EF_DECLINED;
        reads.push_back({0, oq.factor, 0});
        reads.push_back({1, ok_.factor, 0});
    } else {
        for (auto &e : score) {
            if (e.code != EF_EXPR_COLUMN) continue;
            int slot = -1;
            for (auto &r : reads)
                if (r.operand == (int)e.operand && r.column == (int)e.column) slot = r.slot;
            if (slot < 0) {
                if (e.operand > 1) return EF_DECLINED;
                slot = e.operand == 0 ? nq_cols++ : nk_cols++;
                reads.push_back({(int)e.operand, (int)e.column, slot});
            }
            e.column = (uint32_t)slot;
        }
        if (nq_cols == 0 || nk_cols == 0 || nq_cols > 8 || nk_cols > 8) return EF_DECLINED;
    }
    if (ov.factor < 0) return EF_DECLINED;
    if ((!score.empty() && !program_ok(score.data(), score.size())) ||
        (pd->weight_kind == 1 && !program_ok(pd->weight, pd->n_weight))) {
        set_error("a row expression the device can't evaluate");
        return EF_DECLINED;
    }
    if (xs.size() > MAX_DIMS || js.size() > MAX_DIMS || es.size() > MAX_DIMS || ys.size() > MAX_DIMS ||
        outs.size() > MAX_DIMS)
        return EF_DECLINED;

    // Keys and extents (fused.rs:263-300).
    KeyStats *stats;
    EF_TRY(c.alloc(stats, n ? n : 1));
    {
        std::vector<KeyStats> init(n ? n : 1, KeyStats{0, -1, 0, 0});
        EF_TRY(c.upload<KeyStats>(stats, init.data(), init.size()));
    }
    KeyedOperand kq{&q, &oq, {}}, kk{&k, &ok_, {}}, kv{&v, &ov, {}};
    for (KeyedOperand *ko : {&kq, &kk, &kv}) {
        const ef_status st = cast_operand_keys(c, *ko, n, stats);
        if (st != EF_OK) return st;
    }
    std::vector<KeyStats> st(n ? n : 1);
    EF_TRY(c.download<KeyStats>(st.data(), stats, st.size()));
    EF_TRY(c.sync());
    std::vector<uint64_t> extent(n, 0);
    for (uint32_t d = 0; d < n; d++) {
        if (st[d].bad || st[d].lo < 0 || st[d].hi == std::numeric_limits<int64_t>::max()) return EF_DECLINED;
        extent[d] = (uint64_t)(st[d].hi + 1);
    }
    const uint64_t nx = size_of(xs, extent), ne = size_of(es, extent), nj = size_of(js, extent),
                   ny = size_of(ys, extent);
    const uint64_t local_nx = size_of(local_xs, extent), local_nj = size_of(local_js, extent);
    auto dims = [&](const std::vector<uint32_t> &role, const std::vector<int64_t *> &keys) {
        Dims d{};
        d.n = (int)role.size();
        for (size_t i = 0; i < role.size(); i++) {
            d.key[i] = keys[role[i]];
            d.extent[i] = (int64_t)extent[role[i]];
        }
        return d;
    };
    // Placement (fused.rs:624-664): exactly one non-NULL row per cell.
    uint8_t *ones;
    const int64_t max_rows = std::max({q.rows, k.rows, v.rows, (int64_t)1});
    EF_TRY(c.alloc(ones, (size_t)max_rows));
    EF_HIP(hipMemsetAsync(ones, 1, (size_t)max_rows, c.stream));
    PlaceFlags *flags;
    EF_TRY(c.zeros(flags, 1));
    bool any32 = false;
    auto place = [&](const Relation &r, const KeyedOperand &ko, const std::vector<uint32_t> &rd,
                     const std::vector<uint32_t> &cd, int column, double *&m) -> ef_status {
        const uint64_t rows = size_of(rd, extent), cols = size_of(cd, extent);
        const uint64_t cells = mul_sat(rows, cols);
        if ((uint64_t)r.rows != cells || cells >= U32_MAX) return EF_DECLINED;
        if (column < 0 || (size_t)column >= r.columns()) return EF_DECLINED;
        if (!r.batches.empty() && !factor_type_ok(r.types[column], accept32)) return EF_DECLINED;
        for (auto &bt : r.batches)
            if (bt[column].null_count) return EF_DECLINED;
        if (!r.batches.empty() && r.types[column] == Type::Float32) any32 = true;
        double *f;
        uint32_t *present;
        EF_TRY(c.alloc(m, cells));
        EF_TRY(c.alloc(f, cells));
        EF_TRY(c.zeros(present, cells));
        if (cells) EF_TRY(upload_factor(c, r, column, f, nullptr, nullptr));
        return place_b((int64_t)cells, dims(rd, ko.key), dims(cd, ko.key), cols, f, ones, present, m, flags,
                       c.stream);
    };
    double *qm = nullptr, *km = nullptr, *vm = nullptr;
    std::vector<double *> qcols(8, nullptr), kcols(8, nullptr);
    for (const Placement &pl : reads) {
        double *m;
        const ef_status s2 = pl.operand == 0 ? place(q, kq, xs, es, pl.column, m) : place(k, kk, es, js, pl.column, m);
        if (s2 != EF_OK) return s2;
        (pl.operand == 0 ? qcols : kcols)[pl.slot] = m;
    }
    qm = qcols[0];
    km = kcols[0];
    {
        const ef_status s2 = place(v, kv, js, ys, ov.factor, vm);
        if (s2 != EF_OK) return s2;
    }
    PlaceFlags pf;
    EF_TRY(c.download<PlaceFlags>(&pf, flags, 1));
    EF_TRY(c.sync());
    if (pf.duplicate) return EF_DECLINED;
    if (mul_sat(nx, ny) >= U32_MAX) return EF_DECLINED;

    // Each x's allowed keys (fused.rs:350-368), global j.
    std::vector<int64_t> lo(nx), hi(nx);
    const uint64_t lnx = local_nx ? local_nx : 1;
    for (uint64_t x = 0; x < nx; x++) {
        const __int128 xl = (__int128)(x % lnx);
        __int128 l = 0, h = (__int128)local_nj;
        for (const ef_condition &cd : mask) {
            const __int128 bound = xl - (__int128)cd.offset;
            switch (cd.comparison) {
            case EF_GREATER_OR_EQUAL: h = std::min(h, bound + 1); break;
            case EF_GREATER: h = std::min(h, bound); break;
            case EF_LESS_OR_EQUAL: l = std::max(l, bound); break;
            default: l = std::max(l, bound + 1); break;
            }
        }
        h = std::max(h, l);
        auto clamp = [&](__int128 t) { return (int64_t)std::min<__int128>(std::max<__int128>(t, 0), local_nj); };
        const int64_t gs = (int64_t)((x / lnx) * local_nj);
        lo[x] = clamp(l) + gs;
        hi[x] = clamp(h) + gs;
    }
    // Tile pairs (fused.rs:385-397) and those visited (fused.rs:750-770).
    tiles = 0;
    tiles_visited = 0;
    std::vector<uint32_t> diff;
    for (uint64_t x0 = 0; x0 < nx; x0 += 64) {
        const uint64_t x1 = std::min(x0 + 64, nx);
        const uint64_t start = x0 / lnx * local_nj, end = (x1 - 1) / lnx * local_nj + local_nj;
        tiles += (end + 63) / 64 - start / 64;
        int64_t first = (int64_t)nj, last = 0;
        for (uint64_t x = x0; x < x1; x++)
            if (lo[x] < hi[x]) {
                first = std::min(first, lo[x]);
                last = std::max(last, hi[x]);
            }
        if (first >= last) continue;
        const int64_t t0 = first / 64, t1 = (last + 63) / 64;
        diff.assign((size_t)(t1 - t0 + 1), 0);
        for (uint64_t x = x0; x < x1; x++)
            if (lo[x] < hi[x]) {
                diff[(size_t)(lo[x] / 64 - t0)]++;
                diff[(size_t)((hi[x] + 63) / 64 - t0)]--;
            }
        int64_t run = 0;
        for (int64_t t = 0; t < t1 - t0; t++) {
            run += (int32_t)diff[(size_t)t];
            if (run > 0) tiles_visited++;
        }
    }

    // Per-key value statistics (fused.rs:371-384).
    uint8_t *finite, *has_inf;
    double *largest, *smallest, *red;
    int *bad;
    EF_TRY(c.alloc(finite, nj));
    EF_TRY(c.alloc(has_inf, nj));
    EF_TRY(c.alloc(largest, nj));
    EF_TRY(c.alloc(smallest, nj));
    EF_TRY(c.alloc(red, 4));
    EF_TRY(c.zeros(bad, 4));
    EF_TRY(value_stats(vm, (int64_t)nj, (int64_t)ny, finite, has_inf, largest, smallest, c.stream));
    const uint32_t weight_kind = pd->weight_kind;
    const bool standard = prod.row.empty() && weight_kind == 0 && pd->shift == EF_SHIFT_KEPT_MAXIMUM &&
                          pd->normalizer == EF_NORM_KEPT_GROUPS && pd->scale_kind <= EF_SCALE_OVER &&
                          std::isfinite(pd->scale) && (pd->scale_kind != EF_SCALE_OVER || pd->scale != 0.0);
    double *out;
    EF_TRY(c.alloc(out, nx * ny));
    bool fast = false;
    if (standard && nx && nj && ny && ne) {
        // The MFMA path computes the same rows when no value is special and
        // the guards can't trip: every score lies within |scale| |q_x| |k_j|
        // of 0, so every weight is at least lb = exp(-2 |scale| Q K), and
        // with the sum z in [1, terms] Bounds::hold's tests pass for lb.
        const double init[4] = {0.0, std::numeric_limits<double>::infinity(), 0.0, 0.0};
        EF_TRY(c.upload<double>(red, init, 4));
        EF_TRY(reduce_value_stats(finite, largest, smallest, (int64_t)nj, red, bad, c.stream));
        EF_TRY(max_norm(qm, (int64_t)nx, (int64_t)ne, (int64_t)ne, 1, red + 2, bad + 1, c.stream));
        EF_TRY(max_norm(km, (int64_t)nj, (int64_t)ne, 1, (int64_t)nj, red + 3, bad + 2, c.stream));
        double h[4];
        int hb[4];
        EF_TRY(c.download<double>(h, red, 4));
        EF_TRY(c.download<int>(hb, bad, 4));
        EF_TRY(c.sync());
        const double sc = pd->scale_kind == EF_SCALE_TIMES ? std::fabs(pd->scale)
                          : pd->scale_kind == EF_SCALE_OVER ? 1.0 / std::fabs(pd->scale) : 1.0;
        const double lb = std::exp(-2.0 * sc * h[2] * h[3] * (1.0 + 1e-6));
        const double terms = (double)local_nj, normal = 2.0 * DBL_MIN;
        const bool finite_ok = std::isfinite(h[0] * 2.0 * terms);
        const bool no_underflow =
            h[1] == std::numeric_limits<double>::infinity() ||
            (lb * h[1] >= normal && lb / terms >= normal && lb / terms * h[1] >= normal);
        fast = !hb[0] && !hb[1] && !hb[2] && finite_ok && no_underflow && std::isfinite(sc * h[2] * h[3]);
    }
    if (fast) {
        // Batches W, queries local X, keys local J, allowed the same in
        // every batch.
        const uint64_t nw = nx / lnx;
        std::vector<int64_t> spans(2 * lnx);
        for (uint64_t x = 0; x < lnx; x++) {
            spans[2 * x] = lo[x];
            spans[2 * x + 1] = hi[x];
        }
        int64_t *dspans;
        double *kmat;
        EF_TRY(c.alloc(dspans, 2 * lnx));
        EF_TRY(c.upload<int64_t>(dspans, spans.data(), spans.size()));
        EF_TRY(c.alloc(kmat, nj * ne));
        EF_TRY(transpose_f64(km, (int64_t)ne, (int64_t)nj, kmat, c.stream));
        ef_attention_desc d;
        std::memset(&d, 0, sizeof d);
        d.batches = nw;
        d.queries = lnx;
        d.keys = local_nj;
        d.features = ne;
        d.values = ny;
        d.scale = pd->scale_kind == EF_SCALE_TIMES ? pd->scale
                  : pd->scale_kind == EF_SCALE_OVER ? 1.0 / pd->scale : 1.0;
        d.allowed = dspans;
        const uint32_t prec = c.options ? c.options->precision : EF_PRECISION_FP64;
        d.precision = (any32 && prec != EF_PRECISION_FP64) ? prec : EF_PRECISION_FP64;
        double *rm, *rs;   // row statistics: float, or double for FP64
        EF_TRY(c.alloc(rm, nx));
        EF_TRY(c.alloc(rs, nx));
        uint8_t *scratch;
        EF_TRY(c.alloc(scratch, attention_scratch_bytes(d)));
        if (d.precision == EF_PRECISION_FP64) {
            EF_TRY(attention_fwd(d, qm, kmat, vm, out, rm, rs, nullptr, scratch, c.stream));
        } else {
            float *q32, *k32, *v32, *o32;
            EF_TRY(c.alloc(q32, nx * ne));
            EF_TRY(c.alloc(k32, nj * ne));
            EF_TRY(c.alloc(v32, nj * ny));
            EF_TRY(c.alloc(o32, nx * ny));
            EF_TRY(narrow_f32(qm, (int64_t)(nx * ne), q32, c.stream));
            EF_TRY(narrow_f32(kmat, (int64_t)(nj * ne), k32, c.stream));
            EF_TRY(narrow_f32(vm, (int64_t)(nj * ny), v32, c.stream));
            EF_TRY(attention_fwd(d, q32, k32, v32, o32, rm, rs, nullptr, scratch, c.stream));
            EF_TRY(to_f64(Type::Float32, o32, (int64_t)(nx * ny), out, c.stream));
        }
    } else {
        PipeArgs pa{};
        pa.nx = (int64_t)nx;
        pa.ne = (int64_t)ne;
        pa.nj = (int64_t)nj;
        pa.ny = (int64_t)ny;
        pa.q = qm;
        pa.kt = km;
        pa.v = vm;
        for (int i = 0; i < 8; i++) {
            pa.q_cols[i] = qcols[i];
            pa.k_cols[i] = kcols[i];
        }
        pa.n_q_cols = nq_cols;
        pa.n_k_cols = nk_cols;
        pa.sq_q = -1;
        if (!score.empty() && !squared_difference(score, pa)) pa.sq_q = -1;
        if (pd->scale_kind != EF_SCALE_ONE && (!std::isfinite(pd->scale) || pd->scale == 0.0)) pa.sq_mono = 0;
        ef_expr_op *dscore = nullptr, *dweight = nullptr;
        if (!score.empty()) {
            EF_TRY(c.alloc(dscore, score.size()));
            EF_TRY(c.upload<ef_expr_op>(dscore, score.data(), score.size()));
        }
        if (weight_kind == 1) {
            if (!pd->weight || pd->n_weight == 0) {
                set_error("Weight::Of needs its expression");
                return EF_INVALID;
            }
            EF_TRY(c.alloc(dweight, pd->n_weight));
            EF_TRY(c.upload<ef_expr_op>(dweight, pd->weight, pd->n_weight));
        }
        pa.score = dscore;
        pa.n_score = (int)score.size();
        pa.weight = dweight;
        pa.n_weight = weight_kind == 1 ? (int)pd->n_weight : 0;
        pa.scale_kind = pd->scale_kind;
        pa.scale = pd->scale;
        pa.shifted = pd->shift == EF_SHIFT_KEPT_MAXIMUM;
        pa.normalizer = pd->normalizer;
        int64_t *dlo, *dhi;
        EF_TRY(c.alloc(dlo, nx));
        EF_TRY(c.alloc(dhi, nx));
        EF_TRY(c.upload<int64_t>(dlo, lo.data(), nx));
        EF_TRY(c.upload<int64_t>(dhi, hi.data(), nx));
        pa.lo = dlo;
        pa.hi = dhi;
        pa.finite = finite;
        pa.has_inf = has_inf;
        pa.largest = largest;
        pa.smallest = smallest;
        {
            const uint64_t tiles = (nj + 63) / 64;
            uint8_t *tf;
            double *tl, *ts;
            EF_TRY(c.alloc(tf, tiles));
            EF_TRY(c.alloc(tl, tiles));
            EF_TRY(c.alloc(ts, tiles));
            EF_TRY(tile_stats(finite, largest, smallest, (int64_t)nj, tf, tl, ts, c.stream));
            pa.tile_finite = tf;
            pa.tile_largest = tl;
            pa.tile_smallest = ts;
        }
        pa.out = out;
        EF_TRY(c.alloc(pa.special, nx * ny));
        pa.overflowed = bad + 3;
        // Rows and their tiles over the CUs: with few rows, each row's tiles
        // split into runs (parts) so that about four waves run per SIMD,
        // merged in run order; the count depends only on the shape.
        int64_t most_tiles = 0;
        for (uint64_t x = 0; x < nx; x++)
            if (lo[x] < hi[x]) most_tiles = std::max<int64_t>(most_tiles, (hi[x] - 1) / 64 + 1 - lo[x] / 64);
        const int64_t want = 4 * 1024 * 64;
        int parts = (int)std::min<int64_t>(std::min<int64_t>((want + (int64_t)nx - 1) / std::max<int64_t>(nx, 1), 64),
                                           std::max<int64_t>(most_tiles, 1));
        if (parts < 2) parts = 1;
        uint8_t *parts_scratch = nullptr;
        if (parts > 1) EF_TRY(c.alloc(parts_scratch, pipeline_parts_bytes(pa, parts)));
        EF_TRY(pipeline_rows(pa, parts, parts_scratch, c.stream));
        int ov_flag = 0;
        EF_TRY(c.download<int>(&ov_flag, bad + 3, 1));
        EF_TRY(c.sync());
        if (ov_flag) return EF_DECLINED;
    }

    // Rows with allowed keys, every y (fused.rs:435-488).
    std::vector<uint32_t> rows;
    for (uint64_t x = 0; x < nx; x++)
        if (lo[x] < hi[x] && nj > 0) rows.push_back((uint32_t)x);
    const uint64_t nrow = rows.size(), ncell = nrow * ny;
    uint32_t *drows;
    EF_TRY(c.alloc(drows, nrow));
    EF_TRY(c.upload<uint32_t>(drows, rows.data(), nrow));
    OutKeys okeys{};
    okeys.n = (int)outs.size();
    std::vector<uint8_t *> key_out(outs.size());
    for (size_t i = 0; i < outs.size(); i++) {
        const uint32_t d = outs[i];
        const int w = width_of(fields[i].type);
        EF_TRY(c.alloc(key_out[i], ncell * (w ? w : 8)));
        okeys.dst[i] = key_out[i];
        okeys.width[i] = w;
        okeys.in_x[i] = index_in(xs, d) >= 0;
        okeys.stride[i] = stride_in(okeys.in_x[i] ? xs : ys, d, extent);
        okeys.extent[i] = extent[d] ? extent[d] : 1;
    }
    double *vals;
    uint8_t *valid;
    EF_TRY(c.alloc(vals, ncell));
    EF_TRY(c.alloc(valid, ncell));
    EF_TRY(emit((int64_t)ncell, ny, drows, nullptr, nullptr, okeys, out, nullptr, vals, valid, c.stream));
    std::vector<OutColumn> cols;
    for (size_t i = 0; i < outs.size(); i++) {
        OutColumn kc;
        EF_TRY(fetch_column(c, fields[i].type, key_out[i], nullptr, (int64_t)ncell, fields[i].nullable, kc));
        cols.push_back(std::move(kc));
    }
    OutColumn vc;
    EF_TRY(fetch_column(c, Type::Float64, vals, nullptr, (int64_t)ncell, true, vc));
    cols.push_back(std::move(vc));
    result.push_back(std::move(cols));
    return EF_OK;
}

// schema.rs:155-206: the producer's key fields, or v's for the pipeline's
// own dimensions, then a nullable Float64.
bool pipeline_schema(const ef_pipeline_desc *pd, const FoldModel &prod, const Relation &q, const Relation &k,
                     const Relation &v, bool accept32, std::vector<OutField> &fields) {
    std::vector<OutField> pf;
    if (!fold_schema(prod, {&q, &k}, nullptr, accept32, pf)) return false;
    const int vf = pd->values.factor;
    if (vf < 0 || (size_t)vf >= v.columns() || !factor_type_ok(v.types[vf], accept32)) {
        set_error("EinFold factors must be Float64");
        return false;
    }
    fields.clear();
    for (uint32_t i = 0; i < pd->n_outputs; i++) {
        const uint32_t d = pd->outputs[i];
        const int o = prod.output_of(d);
        if (o >= 0) {
            fields.push_back(pf[o]);
            continue;
        }
        int col = -1;
        for (uint32_t j = 0; j < pd->values.n_keys; j++)
            if (pd->values.keys[j].dimension == d) col = (int)pd->values.keys[j].column;
        if (col < 0 || (size_t)col >= v.columns()) {
            set_error("a pipeline output is neither produced nor a key of its values");
            return false;
        }
        fields.push_back(OutField{v.types[col], (bool)v.nullable[col]});
    }
    fields.push_back(OutField{Type::Float64, true});
    return true;
}

}  // namespace ef

using namespace ef;

namespace {

// Reads the relations and computes the schema; EF_OK to go on.
ef_status prepare(const FoldModel &m, const ef_relation *a, const ef_relation *b, const ef_relation *domain,
                  const ef_options *options, Relation &ra, Relation &rb, Relation &rd, bool &has_domain,
                  std::vector<OutField> &fields) {
    if (!read_relation(a, ra) || !read_relation(b, rb)) return EF_INVALID;
    has_domain = domain != nullptr;
    if (has_domain && !read_relation(domain, rd)) return EF_INVALID;
    std::vector<const Relation *> ops = {&ra, &rb};
    if (m.operands.size() != 2) return EF_DECLINED;
    if (!fold_schema(m, ops, has_domain ? &rd : nullptr, options && options->accept_float32, fields))
        return EF_INVALID;
    return EF_OK;
}

// ---------------------------------------------------------------------
// A fold of three operands as two dense folds (chain.rs:55-288)
// ---------------------------------------------------------------------

// chain.rs:57: the largest bound on a partial sum's magnitude taken, 2^1000.
constexpr double CHAIN_LARGEST_BOUND = 1.0715086071862673e301;

// Runs fn(lo, hi, part) over [0, n) in `parts` (1 or POOL_THREADS) parts,
// on the work pool when more than one; the caller combines the parts'
// results in part order.
template <class F> void parallel_parts(int64_t n, int parts, F fn) {
    if (parts <= 1) {
        fn((int64_t)0, n, 0);
        return;
    }
    const int64_t per = (n + parts - 1) / parts;
    const std::function<void(int)> f = [&](int q) {
        const int64_t lo = std::min(per * q, n), hi = std::min(lo + per, n);
        fn(lo, hi, q);
    };
    work_pool().run(f);
}

// One part below 2^18 elements, else the pool's.
int parts_for(int64_t n) { return n >= ((int64_t)1 << 18) ? POOL_THREADS : 1; }

// The largest magnitude of the non-NULL values of a Float64 column as the
// largest of their magnitudes' bit patterns (for finite doubles the same
// order), which is at least 0x7ff0... exactly when one is inf or NaN.
uint64_t largest_magnitude_bits(const ColumnView &v) {
    const uint64_t *x = reinterpret_cast<const uint64_t *>(v.values);
    constexpr uint64_t ABS = 0x7fffffffffffffffULL;
    const int parts = parts_for(v.length);
    uint64_t part_max[POOL_THREADS] = {};
    parallel_parts(v.length, parts, [&](int64_t lo, int64_t hi, int q) {
        uint64_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
        int64_t i = lo;
        if (v.null_count == 0) {
            for (; i + 4 <= hi; i += 4) {
                const uint64_t a = x[i] & ABS, b = x[i + 1] & ABS, c = x[i + 2] & ABS, d = x[i + 3] & ABS;
                m0 = a > m0 ? a : m0;
                m1 = b > m1 ? b : m1;
                m2 = c > m2 ? c : m2;
                m3 = d > m3 ? d : m3;
            }
        }
        for (; i < hi; i++) {
            const uint64_t a = v.valid(i) ? x[i] & ABS : 0;
            m0 = a > m0 ? a : m0;
        }
        m0 = std::max(std::max(m0, m1), std::max(m2, m3));
        part_max[q] = m0;
    });
    uint64_t m = 0;
    for (int q = 0; q < parts; q++) m = std::max(m, part_max[q]);
    return m;
}

// chain.rs:155-181: every non-NULL factor is a finite Float64, and the
// product over the operands of their largest magnitude times their rows,
// which bounds every partial sum in either order, is at most 2^1000.
bool chain_guarded(const FoldModel &m, const std::vector<const Relation *> &ops) {
    constexpr uint64_t INF_BITS = 0x7ff0000000000000ULL;
    double bound = 1.0;
    for (size_t o = 0; o < ops.size(); o++) {
        const int c = m.operands[o].factor;
        const Relation &r = *ops[o];
        if (c < 0 || r.types[c] != Type::Float64) return false;
        double largest = 0.0;
        int64_t rows = 0;
        for (size_t b = 0; b < r.batches.size(); b++) {
            rows += r.batch_rows[b];
            const uint64_t bits = largest_magnitude_bits(r.batches[b][c]);
            if (bits >= INF_BITS) return false;
            double big;
            std::memcpy(&big, &bits, 8);
            largest = std::max(largest, big);
        }
        bound *= largest * (double)rows;
    }
    return bound <= CHAIN_LARGEST_BOUND;
}

// The smallest and largest key of a column of integer type T (as i64; an
// unsigned 64-bit key of 2^63 or more comes out negative: too large for
// i64), over parallel parts with four accumulators each.
template <class T> void key_range(const T *k, int64_t n, int64_t &lo_out, int64_t &hi_out) {
    const int parts = parts_for(n);
    int64_t part_lo[POOL_THREADS], part_hi[POOL_THREADS];
    parallel_parts(n, parts, [&](int64_t lo, int64_t hi, int q) {
        int64_t l0 = INT64_MAX, l1 = INT64_MAX, h0 = INT64_MIN, h1 = INT64_MIN;
        int64_t i = lo;
        for (; i + 2 <= hi; i += 2) {
            const int64_t a = (int64_t)k[i], b = (int64_t)k[i + 1];
            l0 = a < l0 ? a : l0;
            h0 = a > h0 ? a : h0;
            l1 = b < l1 ? b : l1;
            h1 = b > h1 ? b : h1;
        }
        for (; i < hi; i++) {
            const int64_t a = (int64_t)k[i];
            l0 = a < l0 ? a : l0;
            h0 = a > h0 ? a : h0;
        }
        part_lo[q] = std::min(l0, l1);
        part_hi[q] = std::max(h0, h1);
    });
    int64_t l = INT64_MAX, h = INT64_MIN;
    for (int q = 0; q < parts; q++) {
        l = std::min(l, part_lo[q]);
        h = std::max(h, part_hi[q]);
    }
    lo_out = l;
    hi_out = h;
}

// dense.rs:97-111 for one batch's key column: integers without NULLs, each
// in [0, 2^63) and below i64::MAX; the extent is the largest plus one.
bool chain_key_extent(const ColumnView &v, int64_t &extent) {
    if (!is_integer(v.type) || v.null_count > 0) return false;
    if (v.length == 0) {
        extent = 0;
        return true;
    }
    int64_t lo = 0, hi = -1;
    switch (v.type) {
    case Type::Int8: key_range(reinterpret_cast<const int8_t *>(v.values), v.length, lo, hi); break;
    case Type::UInt8: key_range(reinterpret_cast<const uint8_t *>(v.values), v.length, lo, hi); break;
    case Type::Int16: key_range(reinterpret_cast<const int16_t *>(v.values), v.length, lo, hi); break;
    case Type::UInt16: key_range(reinterpret_cast<const uint16_t *>(v.values), v.length, lo, hi); break;
    case Type::Int32: key_range(reinterpret_cast<const int32_t *>(v.values), v.length, lo, hi); break;
    case Type::UInt32: key_range(reinterpret_cast<const uint32_t *>(v.values), v.length, lo, hi); break;
    case Type::Int64: key_range(reinterpret_cast<const int64_t *>(v.values), v.length, lo, hi); break;
    default: key_range(reinterpret_cast<const uint64_t *>(v.values), v.length, lo, hi); break;
    }
    if (lo < 0 || hi == std::numeric_limits<int64_t>::max()) return false;
    extent = hi + 1;
    return true;
}

// chain.rs:197-221: an operand's dimensions in order; the union of two
// sorted lists.
std::vector<uint32_t> chain_dims(const OperandModel &op) {
    std::vector<uint32_t> d;
    for (auto &k : op.keys) d.push_back(k.first);
    std::sort(d.begin(), d.end());
    return d;
}
std::vector<uint32_t> chain_union(const std::vector<uint32_t> &a, const std::vector<uint32_t> &b) {
    std::vector<uint32_t> u(a);
    u.insert(u.end(), b.begin(), b.end());
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    return u;
}
int chain_pos(const std::vector<uint32_t> &local, uint32_t d) {
    auto it = std::find(local.begin(), local.end(), d);
    return it == local.end() ? -1 : (int)(it - local.begin());
}
// The product of the dimensions' extents; false past 128 bits.
bool chain_cells(const std::vector<uint32_t> &dims, const std::vector<int64_t> &extent, unsigned __int128 &n) {
    n = 1;
    for (uint32_t d : dims) {
        const unsigned __int128 e = (unsigned __int128)extent[d];
        if (e != 0 && n > ~(unsigned __int128)0 / e) return false;
        n *= e;
    }
    return true;
}
// chain.rs:223-231: of the first pair's dimensions, those the first fold
// keeps: the outputs and those operand `rest` joins on.
std::vector<uint32_t> chain_kept(const FoldModel &m, const std::vector<uint32_t> &pair, size_t rest) {
    std::vector<uint32_t> kept;
    for (uint32_t d : pair)
        if (m.output_of(d) >= 0 || m.operands[rest].column_of(d) >= 0) kept.push_back(d);
    return kept;
}
// A copy of `op` whose keys name each dimension by its index in `local`.
OperandModel chain_renumber(const OperandModel &op, const std::vector<uint32_t> &local) {
    OperandModel r;
    r.factor = op.factor;
    for (auto &k : op.keys) r.keys.emplace_back((uint32_t)chain_pos(local, k.first), k.second);
    return r;
}

// chain.rs:233-288: the first fold, of operands x and y, keeping one key
// column per kept dimension, in order, then the sum; the second, of that
// result and operand `rest`, with the chain's outputs.
void chain_steps(const FoldModel &m, size_t x, size_t y, size_t rest, FoldModel &first, FoldModel &second) {
    const std::vector<uint32_t> pair = chain_union(chain_dims(m.operands[x]), chain_dims(m.operands[y]));
    const std::vector<uint32_t> kept = chain_kept(m, pair, rest);
    first = FoldModel();
    for (uint32_t d : pair) first.equality.push_back(m.equality[d]);
    first.operands = {chain_renumber(m.operands[x], pair), chain_renumber(m.operands[y], pair)};
    for (uint32_t d : kept) first.outputs.push_back(ef_output{(uint32_t)chain_pos(pair, d), -1, 0});
    const std::vector<uint32_t> local = chain_union(kept, chain_dims(m.operands[rest]));
    second = FoldModel();
    for (uint32_t d : local) second.equality.push_back(m.equality[d]);
    OperandModel t;
    for (size_t c = 0; c < kept.size(); c++) t.keys.emplace_back((uint32_t)chain_pos(local, kept[c]), (uint32_t)c);
    t.factor = (int)kept.size();
    second.operands = {t, chain_renumber(m.operands[rest], local)};
    for (auto &o : m.outputs) second.outputs.push_back(ef_output{(uint32_t)chain_pos(local, o.dimension), -1, 0});
}

// chain.rs:125-153: the dense fold of a and b, or, where it declines, of b
// and a (the dense algorithm takes holes in its second operand only).
ef_status chain_either(const FoldModel &f, const Relation &a, const Relation &b, const std::vector<OutField> &fields,
                       const ef_options *options, std::vector<std::vector<OutColumn>> &out) {
    {
        Call c(options);
        const ef_status st = fold_dense_impl(f, a, b, nullptr, fields, c, out, nullptr);
        (void)hipStreamSynchronize(c.stream);
        if (st != EF_DECLINED) return st;
    }
    FoldModel swapped = f;
    std::swap(swapped.operands[0], swapped.operands[1]);
    out.clear();
    Call c(options);
    const ef_status st = fold_dense_impl(swapped, b, a, nullptr, fields, c, out, nullptr);
    (void)hipStreamSynchronize(c.stream);
    return st;
}

// The first fold's output batches read as a relation (no copy).
Relation chain_relation(const std::vector<std::vector<OutColumn>> &batches, const std::vector<OutField> &fields) {
    Relation r;
    for (const OutField &f : fields) {
        r.types.push_back(f.type);
        r.nullable.push_back(f.nullable);
        r.formats.push_back(format_of(f.type));
    }
    for (const auto &cols : batches) {
        std::vector<ColumnView> views(cols.size());
        for (size_t c = 0; c < cols.size(); c++) {
            ColumnView &v = views[c];
            v.type = cols[c].type;
            v.values = cols[c].values ? cols[c].values->data : nullptr;
            v.validity = cols[c].validity ? cols[c].validity->data : nullptr;
            v.length = cols[c].length;
            v.null_count = cols[c].null_count;
        }
        const int64_t n = cols.empty() ? 0 : cols[0].length;
        r.batches.push_back(std::move(views));
        r.batch_rows.push_back(n);
        r.rows += n;
    }
    return r;
}

// chain.rs:64-123.
ef_status fold_chain_impl(const FoldModel &m, const std::vector<const Relation *> &ops,
                          const std::vector<OutField> &fields, const ef_options *options,
                          std::vector<std::vector<OutColumn>> &out) {
    if (!chain_guarded(m, ops)) return EF_DECLINED;
    // chain.rs:186-198: each dimension's extent over every operand's keys.
    std::vector<int64_t> extent(m.equality.size(), 0);
    for (size_t o = 0; o < ops.size(); o++)
        for (const auto &cols : ops[o]->batches)
            for (auto &k : m.operands[o].keys) {
                int64_t e;
                if (!chain_key_extent(cols[k.second], e)) return EF_DECLINED;
                extent[k.first] = std::max(extent[k.first], e);
            }
    // chain.rs:80-107: the pair contracted first, the cheapest of those that
    // share a dimension (cells of the pair plus cells of the second fold),
    // the first of equals.
    const size_t cand[3][3] = {{0, 1, 2}, {1, 2, 0}, {0, 2, 1}};
    int best = -1;
    unsigned __int128 best_cost = 0;
    for (int i = 0; i < 3; i++) {
        const size_t x = cand[i][0], y = cand[i][1], rest = cand[i][2];
        const std::vector<uint32_t> dx = chain_dims(m.operands[x]), dy = chain_dims(m.operands[y]);
        bool shares = false;
        for (uint32_t d : dx) shares |= std::find(dy.begin(), dy.end(), d) != dy.end();
        if (!shares) continue;
        const std::vector<uint32_t> pair = chain_union(dx, dy);
        const std::vector<uint32_t> second = chain_union(chain_kept(m, pair, rest), chain_dims(m.operands[rest]));
        unsigned __int128 c1, c2;
        if (!chain_cells(pair, extent, c1) || !chain_cells(second, extent, c2) || c1 + c2 < c1) continue;
        if (best < 0 || c1 + c2 < best_cost) {
            best = i;
            best_cost = c1 + c2;
        }
    }
    if (best < 0) return EF_DECLINED;
    const size_t x = cand[best][0], y = cand[best][1], rest = cand[best][2];
    FoldModel first, second;
    chain_steps(m, x, y, rest, first, second);
    const bool accept32 = options && options->accept_float32;
    std::vector<OutField> first_fields;
    if (!fold_schema(first, {ops[x], ops[y]}, nullptr, accept32, first_fields)) return EF_INVALID;
    std::vector<std::vector<OutColumn>> t;
    EF_TRY(chain_either(first, *ops[x], *ops[y], first_fields, options, t));
    settle(t);
    const Relation rt = chain_relation(t, first_fields);
    return chain_either(second, rt, *ops[rest], fields, options, out);
}

// Runs `body` catching host allocation failures, so no exception crosses
// the ABI.
template <class F> ef_status guarded(F body) {
    try {
        return body();
    } catch (const std::bad_alloc &) {
        set_error("host memory exhausted");
        return EF_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        set_error(e.what());
        return EF_DEVICE_ERROR;
    }
}

}  // namespace

extern "C" {

void ef_result_release(ef_result *r) {
    if (!r) return;
    for (size_t i = 0; i < r->n_batches; i++)
        if (r->batches[i].release) r->batches[i].release(&r->batches[i]);
    delete[] r->batches;
    r->batches = nullptr;
    r->n_batches = 0;
    if (r->schema.release) r->schema.release(&r->schema);
}

ef_status ef_fold_dense(const ef_fold_desc *fold, const ef_relation *a, const ef_relation *b,
                        const ef_relation *domain, const ef_options *options, ef_result *out) {
    clear_error();
    if (!out) {
        set_error("no result to write");
        return EF_INVALID;
    }
    std::memset(out, 0, sizeof *out);
    return guarded([&]() -> ef_status {
        FoldModel m;
        if (!parse_fold(fold, m)) return EF_INVALID;
        Relation ra, rb, rd;
        bool has_domain;
        std::vector<OutField> fields;
        EF_TRY(prepare(m, a, b, domain, options, ra, rb, rd, has_domain, fields));
        Call c(options);
        std::vector<std::vector<OutColumn>> batches;
        ef_status st = fold_dense_impl(m, ra, rb, has_domain ? &rd : nullptr, fields, c, batches, nullptr);
        if (st != EF_OK) {
            (void)hipStreamSynchronize(c.stream);
            return st;
        }
        EF_TRY(c.sync());
        put_result(std::move(batches), fields, EF_ALGORITHM_DENSE, out);
        return EF_OK;
    });
}

ef_status ef_fold_dense_many(size_t count, const ef_fold_desc *const *folds, const ef_relation *a,
                             const ef_relation *b, const ef_relation *const *domains, const ef_options *options,
                             ef_result *out, ef_status *status) {
    clear_error();
    if (count && (!folds || !a || !b || !out || !status)) {
        set_error("folds, operands, results and statuses are required");
        return EF_INVALID;
    }
    for (size_t i = 0; i < count; i++) {
        std::memset(&out[i], 0, sizeof out[i]);
        status[i] = EF_DECLINED;
    }
    return guarded([&]() -> ef_status {
        // Every fold is read, its keys cast and its operands placed; the
        // MFMA contractions wait in one batch, then run as one launch; then
        // each deferred fold widens its sums and emits.
        struct One {
            FoldModel m;
            Relation ra, rb, rd;
            bool has_domain = false;
            std::vector<OutField> fields;
            std::vector<std::vector<OutColumn>> batches;
            std::function<ef_status()> finish;
        };
        std::deque<One> ones(count);
        Call c(options);
        DenseBatch batch;
        for (size_t i = 0; i < count; i++) {
            One &o = ones[i];
            if (!parse_fold(folds[i], o.m)) {
                status[i] = EF_INVALID;
                continue;
            }
            ef_status st = prepare(o.m, &a[i], &b[i], domains ? domains[i] : nullptr, options, o.ra, o.rb, o.rd,
                                   o.has_domain, o.fields);
            if (st == EF_OK) {
                batch.finish = nullptr;
                st = fold_dense_impl(o.m, o.ra, o.rb, o.has_domain ? &o.rd : nullptr, o.fields, c, o.batches, &batch);
                o.finish = std::move(batch.finish);
                batch.finish = nullptr;
            }
            if (st == EF_DEVICE_ERROR) {
                (void)hipStreamSynchronize(c.stream);
                return st;   // a device failure ends the call
            }
            status[i] = st;
        }
        if (!batch.problems.empty())
            EF_TRY(grouped_run_once(batch.problems.data(), batch.problems.size(), batch.precision,
                                    batch.density.data(), c.stream));
        for (size_t i = 0; i < count; i++) {
            One &o = ones[i];
            if (status[i] != EF_OK || !o.finish) continue;
            const ef_status st = o.finish();
            if (st != EF_OK) {
                (void)hipStreamSynchronize(c.stream);
                return st;
            }
        }
        EF_HIP(hipStreamSynchronize(c.stream));
        for (size_t i = 0; i < count; i++)
            if (status[i] == EF_OK) put_result(std::move(ones[i].batches), ones[i].fields, EF_ALGORITHM_DENSE, &out[i]);
        return EF_OK;
    });
}

ef_status ef_fold(const ef_fold_desc *fold, const ef_relation *a, const ef_relation *b, const ef_relation *domain,
                  const ef_options *options, ef_result *out) {
    clear_error();
    if (!out) {
        set_error("no result to write");
        return EF_INVALID;
    }
    std::memset(out, 0, sizeof *out);
    return guarded([&]() -> ef_status {
        FoldModel m;
        if (!parse_fold(fold, m)) return EF_INVALID;
        if (m.operands.size() != 2) {
            set_error("a fold of two operands");
            return EF_INVALID;
        }
        if (options && options->domain_rows == EF_DOMAIN_ROWS_UNIQUE_KEYS) {
            set_error("only fold_dense checks that a domain's keys are unique");
            return EF_INVALID;
        }
        Relation ra, rb, rd;
        bool has_domain;
        std::vector<OutField> fields;
        EF_TRY(prepare(m, a, b, domain, options, ra, rb, rd, has_domain, fields));
        std::vector<std::vector<OutColumn>> batches;
        {
            Call c(options);
            ef_status st = fold_dense_impl(m, ra, rb, has_domain ? &rd : nullptr, fields, c, batches, nullptr);
            (void)hipStreamSynchronize(c.stream);
            if (st == EF_OK) {
                put_result(std::move(batches), fields, EF_ALGORITHM_DENSE, out);
                return EF_OK;
            }
            if (st != EF_DECLINED) return st;
        }
        // algorithm.rs:114-133: the sorting algorithm computes no masks, and
        // takes each operand as one batch.
        if (!m.conditions.empty()) {
            set_error("the discrimination algorithm computes no masks");
            return EF_INVALID;
        }
        Call c(options);
        EF_TRY(c.reserve((double)relation_bytes(ra) + (double)relation_bytes(rb), "the operands as one batch each"));
        batches.clear();
        ef_status st = fold_discriminate_impl(m, ra, rb, has_domain ? &rd : nullptr, fields, c, batches);
        if (st != EF_OK) {
            (void)hipStreamSynchronize(c.stream);
            return st;
        }
        EF_TRY(c.sync());
        put_result(std::move(batches), fields, EF_ALGORITHM_DISCRIMINATE, out);
        return EF_OK;
    });
}

ef_status ef_fold_chain(const ef_fold_desc *fold, const ef_relation *a, const ef_relation *b, const ef_relation *c,
                        const ef_options *options, ef_result *out) {
    clear_error();
    if (!out) {
        set_error("no result to write");
        return EF_INVALID;
    }
    std::memset(out, 0, sizeof *out);
    return guarded([&]() -> ef_status {
        FoldModel m;
        if (!parse_fold(fold, m)) return EF_INVALID;
        // chain.rs:70-78: a SUM of the product of three factors, no domain,
        // no conditions.
        if (m.operands.size() != 3 || !m.row.empty() || m.has_domain || !m.conditions.empty()) return EF_DECLINED;
        for (auto &op : m.operands)
            if (op.factor < 0) return EF_DECLINED;
        Relation ra, rb, rc;
        if (!read_relation(a, ra) || !read_relation(b, rb) || !read_relation(c, rc)) return EF_INVALID;
        const std::vector<const Relation *> ops = {&ra, &rb, &rc};
        std::vector<OutField> fields;
        if (!fold_schema(m, ops, nullptr, options && options->accept_float32, fields)) return EF_INVALID;
        std::vector<std::vector<OutColumn>> batches;
        EF_TRY(fold_chain_impl(m, ops, fields, options, batches));
        put_result(std::move(batches), fields, EF_ALGORITHM_DENSE, out);
        return EF_OK;
    });
}

ef_status ef_pipeline(const ef_pipeline_desc *pipeline, const ef_relation *q, const ef_relation *k,
                      const ef_relation *v, const ef_options *options, ef_result *out) {
    clear_error();
    if (!out || !pipeline) {
        set_error("a pipeline and a result are required");
        return EF_INVALID;
    }
    std::memset(out, 0, sizeof *out);
    return guarded([&]() -> ef_status {
        FoldModel prod;
        if (!parse_fold(&pipeline->producer, prod)) return EF_INVALID;
        const uint32_t n = (uint32_t)prod.equality.size() + pipeline->n_dimensions;
        for (uint32_t i = 0; i < pipeline->n_outputs; i++)
            if (pipeline->outputs[i] >= n) {
                set_error("a pipeline output names a dimension it doesn't have");
                return EF_INVALID;
            }
        for (uint32_t i = 0; i < pipeline->values.n_keys; i++)
            if (pipeline->values.keys[i].dimension >= n) {
                set_error("a value key names a dimension the pipeline doesn't have");
                return EF_INVALID;
            }
        if (pipeline->weight_kind > 1 || pipeline->normalizer > EF_NORM_NONE || pipeline->scale_kind > EF_SCALE_OVER) {
            set_error("unknown weight, scale or normalizer");
            return EF_INVALID;
        }
        Relation rq, rk, rv;
        if (!read_relation(q, rq) || !read_relation(k, rk) || !read_relation(v, rv)) return EF_INVALID;
        std::vector<OutField> fields;
        if (prod.operands.size() != 2) return EF_DECLINED;
        if (!pipeline_schema(pipeline, prod, rq, rk, rv, options && options->accept_float32, fields))
            return EF_INVALID;
        Call c(options);
        std::vector<std::vector<OutColumn>> batches;
        uint64_t visited = 0, tiles = 0;
        ef_status st = pipeline_impl(pipeline, prod, rq, rk, rv, fields, c, batches, visited, tiles);
        if (st != EF_OK) {
            (void)hipStreamSynchronize(c.stream);
            return st;
        }
        EF_TRY(c.sync());
        put_result(std::move(batches), fields, EF_ALGORITHM_FUSED, out);
        out->tiles_visited = visited;
        out->tiles = tiles;
        return EF_OK;
    });
}

}  // extern "C"
