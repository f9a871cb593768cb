// Device helpers of the operator path (ops.hip), callable from the host.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "arrow_io.h"
#include "common.h"

namespace ef {

constexpr int MAX_DIMS = 8;
constexpr int MAX_CONDS = 8;

// A key column's bounds as dense.rs:97-111 computes them: lo starts at 0
// (so a negative key shows), hi at -1 (an empty column has extent 0); bad
// marks an unsigned key of 2^63 or more.
struct KeyStats {
    int64_t lo;
    int64_t hi;
    int bad;
    int pad;
};

// Dimensions of one role, first slowest: each key column and its extent.
struct Dims {
    const int64_t *key[MAX_DIMS];
    int64_t extent[MAX_DIMS];
    int n;
};

struct PlaceFlags {
    int duplicate;
    int not_finite;      // a's valued factors
    int nulls;
    int b_not_finite;    // b's valued factors
};

// How a domain's key columns locate a cell of the dense output.
struct CellKeys {
    const int64_t *key[MAX_DIMS];
    const uint8_t *valid[MAX_DIMS];
    int64_t extent[MAX_DIMS];
    int64_t stride[MAX_DIMS];
    int n;
};

// Conditions between output keys, each side a digit of x's or y's position.
struct Conds {
    int op[MAX_CONDS];
    int left_in_x[MAX_CONDS], right_in_x[MAX_CONDS];
    uint64_t left_stride[MAX_CONDS], left_extent[MAX_CONDS];
    uint64_t right_stride[MAX_CONDS], right_extent[MAX_CONDS];
    int64_t offset[MAX_CONDS];
    int n;
};

// The output's key columns: destination, width in bytes, and the digit.
struct OutKeys {
    void *dst[MAX_DIMS];
    int width[MAX_DIMS];
    int in_x[MAX_DIMS];
    uint64_t stride[MAX_DIMS], extent[MAX_DIMS];
    int n;
};

ef_status scan_u32(const uint32_t *in, uint32_t *out, size_t n, void *tmp, hipStream_t s);
size_t scan_scratch_bytes(size_t n);

ef_status cast_keys(Type t, const void *in, int64_t n, int64_t *out, KeyStats *stats, hipStream_t s);
ef_status to_f64(Type t, const void *in, int64_t n, double *out, hipStream_t s);
ef_status place_a(int64_t n, const Dims &xs, const Dims &zs, uint64_t nz, const double *f, const uint8_t *fvalid,
                  uint32_t *held, double *a, uint8_t *reached, uint8_t *valued, PlaceFlags *flags, hipStream_t s);
ef_status place_b(int64_t n, const Dims &wzs, const Dims &ys, uint64_t ny, const double *f, const uint8_t *fvalid,
                  uint32_t *present, double *b, PlaceFlags *flags, hipStream_t s);
ef_status locate(int64_t n, const CellKeys &ck, int64_t *cell, hipStream_t s);
ef_status unique_cells(int64_t n, const int64_t *cell, uint32_t *seen, int *repeated, hipStream_t s);
ef_status reached_rows(int64_t nx, const uint8_t *reached, uint32_t *tmp32, uint32_t *excl, uint32_t *rows,
                       void *scan_tmp, hipStream_t s);
ef_status keep_cells(int64_t n, uint64_t ny, const uint32_t *rows, const Conds &c, uint32_t *keep, hipStream_t s);
ef_status emit(int64_t n, uint64_t ny, const uint32_t *rows, const uint32_t *keep, const uint32_t *excl,
               const OutKeys &ok, const double *out, const uint8_t *valued, double *values, uint8_t *valid,
               hipStream_t s);
ef_status domain_values(int64_t n, const int64_t *cell, uint64_t ny, const double *out, const uint8_t *valued,
                        const Conds &c, int has_missing, double missing, double *values, uint8_t *valid,
                        hipStream_t s);
ef_status unpack_bits(int64_t n, const uint8_t *bits, int64_t offset, uint8_t *out, hipStream_t s);
ef_status widen_keys(Type t, int64_t n, const void *in, uint8_t *valid, int64_t *out, hipStream_t s);
ef_status pack_bits(int64_t n, const uint8_t *bytes, uint8_t *bits, unsigned long long *nulls, hipStream_t s);

}  // namespace ef
