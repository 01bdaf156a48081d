// Datatypes and reduction operators.
//
// Sizes and enum order: gpu_libs/gpu_mpi/mpi_types_list.cuh:3-33 and
// datatypes.cu:10-15.  Operator semantics and the type x op legality table:
// gpu_libs/gpu_mpi/operators.cu:73-200, with MPI_LXOR as the MPI logical
// exclusive or.  The fold order is the reference's generic Reduce
// (mpi.cu:209-225, operators.cu:216-218): acc = x_0, then acc = op(x_r, acc)
// for r = 1..P-1 -- deterministic and identical for every datatype.
#pragma once

#include "device.hpp"

namespace gpumpi {
namespace rt {

__host__ __device__ inline int basic_type_size(int dt) {
    switch (dt) {
    case MPI_CHAR: case MPI_SIGNED_CHAR: case MPI_UNSIGNED_CHAR: case MPI_C_BOOL:
    case MPI_INT8_T: case MPI_UINT8_T: case MPI_BYTE: return 1;
    case MPI_SHORT: case MPI_UNSIGNED_SHORT: case MPI_INT16_T: case MPI_UINT16_T: return 2;
    case MPI_INT: case MPI_UNSIGNED: case MPI_FLOAT: case MPI_WCHAR: case MPI_INT32_T:
    case MPI_UINT32_T: return 4;
    case MPI_LONG: case MPI_LONG_LONG_INT: case MPI_LONG_LONG: case MPI_UNSIGNED_LONG:
    case MPI_UNSIGNED_LONG_LONG: case MPI_DOUBLE: case MPI_LONG_DOUBLE: case MPI_INT64_T:
    case MPI_UINT64_T: case MPI_C_COMPLEX: case MPI_C_FLOAT_COMPLEX: return 8;
    case MPI_C_DOUBLE_COMPLEX: case MPI_C_LONG_DOUBLE_COMPLEX: return 16;
    default: return -1;
    }
}

// Contiguous derived types are their own handles:
// GPUMPI_CONTIGUOUS_FLAG | count << 5 | basic type.
constexpr int kContigShift = 5;
constexpr int kContigMaxCount = (1 << 25) - 1;
__host__ __device__ inline bool type_is_contig(int dt) { return dt > 0 && (dt & GPUMPI_CONTIGUOUS_FLAG) != 0; }
__host__ __device__ inline int type_base(int dt) { return type_is_contig(dt) ? (dt & 31) : dt; }
__host__ __device__ inline int type_count(int dt) {
    return type_is_contig(dt) ? ((dt >> kContigShift) & kContigMaxCount) : 1;
}
// Bytes of one element of dt, or -1 for no valid datatype.
__host__ __device__ inline int type_size(int dt) {
    if (!type_is_contig(dt)) return basic_type_size(dt);
    const int b = basic_type_size(dt & 31);
    const int64_t n = (int64_t)type_count(dt) * b;
    return b > 0 && n > 0 && n <= 0x7fffffff ? (int)n : -1;
}

enum TypeClass { TC_NONE, TC_INT, TC_FLOAT, TC_LOGICAL, TC_COMPLEX, TC_BYTE, TC_TEXT };
__host__ __device__ inline int type_class(int dt) {
    switch (dt) {
    case MPI_SHORT: case MPI_INT: case MPI_LONG: case MPI_LONG_LONG_INT: case MPI_LONG_LONG:
    case MPI_SIGNED_CHAR: case MPI_UNSIGNED_CHAR: case MPI_UNSIGNED_SHORT: case MPI_UNSIGNED:
    case MPI_UNSIGNED_LONG: case MPI_UNSIGNED_LONG_LONG: case MPI_INT8_T: case MPI_INT16_T:
    case MPI_INT32_T: case MPI_INT64_T: case MPI_UINT8_T: case MPI_UINT16_T: case MPI_UINT32_T:
    case MPI_UINT64_T: return TC_INT;
    case MPI_FLOAT: case MPI_DOUBLE: case MPI_LONG_DOUBLE: return TC_FLOAT;
    case MPI_C_BOOL: return TC_LOGICAL;
    case MPI_C_COMPLEX: case MPI_C_FLOAT_COMPLEX: case MPI_C_DOUBLE_COMPLEX:
    case MPI_C_LONG_DOUBLE_COMPLEX: return TC_COMPLEX;
    case MPI_BYTE: return TC_BYTE;
    case MPI_CHAR: case MPI_WCHAR: return TC_TEXT;
    default: return TC_NONE;
    }
}

enum BuiltinOp { OP_MAX, OP_MIN, OP_SUM, OP_PROD, OP_LAND, OP_BAND, OP_LOR, OP_BOR, OP_LXOR, OP_BXOR,
                 OP_MAXLOC, OP_MINLOC };

// operators.cu:153-200.  MAXLOC/MINLOC have no legal basic datatype.
__host__ __device__ inline bool op_allowed(int dt, int op) {
    const int c = type_class(dt);
    switch (op) {
    case OP_MAX: case OP_MIN: return c == TC_INT || c == TC_FLOAT;
    case OP_SUM: case OP_PROD: return c == TC_INT || c == TC_FLOAT || c == TC_COMPLEX;
    case OP_LAND: case OP_LOR: case OP_LXOR: return c == TC_INT || c == TC_LOGICAL;
    case OP_BAND: case OP_BOR: case OP_BXOR: return c == TC_INT || c == TC_BYTE;
    default: return false;
    }
}

__device__ __forceinline__ bool type_signed_int(int dt) {
    switch (dt) {
    case MPI_SHORT: case MPI_INT: case MPI_LONG: case MPI_LONG_LONG_INT: case MPI_LONG_LONG:
    case MPI_SIGNED_CHAR: case MPI_INT8_T: case MPI_INT16_T: case MPI_INT32_T: case MPI_INT64_T:
        return true;
    default: return false;
    }
}

struct cfloat2 { float re, im; };
struct cdouble2 { double re, im; };

// Functors: a = incoming x_r, b = accumulator (operators.cu:73-121).
struct FMax { template <class T> __device__ __forceinline__ T operator()(T a, T b) const { return a > b ? a : b; } };
struct FMin { template <class T> __device__ __forceinline__ T operator()(T a, T b) const { return a > b ? b : a; } };
struct FSum {
    template <class T> __device__ __forceinline__ T operator()(T a, T b) const { return a + b; }
    __device__ __forceinline__ cfloat2 operator()(cfloat2 a, cfloat2 b) const { return {a.re + b.re, a.im + b.im}; }
    __device__ __forceinline__ cdouble2 operator()(cdouble2 a, cdouble2 b) const { return {a.re + b.re, a.im + b.im}; }
};
struct FProd {
    // 8- and 16-bit operands would be promoted to (signed) int, whose
    // overflow is undefined; multiply them as unsigned int and truncate.
    template <class T> __device__ __forceinline__ T operator()(T a, T b) const {
        if constexpr (sizeof(T) < sizeof(unsigned) && T(-1) > T(0)) return (T)((unsigned)a * (unsigned)b);
        else return a * b;
    }
    __device__ __forceinline__ cfloat2 operator()(cfloat2 a, cfloat2 b) const {
#pragma clang fp contract(off)
        return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    }
    __device__ __forceinline__ cdouble2 operator()(cdouble2 a, cdouble2 b) const {
#pragma clang fp contract(off)
        return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    }
};
struct FLAnd { template <class T> __device__ __forceinline__ T operator()(T a, T b) const { return (T)((a != 0) && (b != 0)); } };
struct FLOr { template <class T> __device__ __forceinline__ T operator()(T a, T b) const { return (T)((a != 0) || (b != 0)); } };
struct FLXor { template <class T> __device__ __forceinline__ T operator()(T a, T b) const { return (T)((a != 0) != (b != 0)); } };
struct FBAnd { template <class T> __device__ __forceinline__ T operator()(T a, T b) const { return (T)(a & b); } };
struct FBOr { template <class T> __device__ __forceinline__ T operator()(T a, T b) const { return (T)(a | b); } };
struct FBXor { template <class T> __device__ __forceinline__ T operator()(T a, T b) const { return (T)(a ^ b); } };

// Fold of the members' inputs over elements [lo, hi).
//   ins:  address of member r's input = ins_w[r * stride_words]
//   outs: written to member `root`'s output only (root >= 0) or to every
//         member's output (root < 0), address = outs_w[r * stride_words]
// Worker t of nt; `vec` says every pointer is 16-byte aligned and lo is a
// multiple of the vector width.
struct FoldArgs {
    const uint64_t* ins_w;
    const uint64_t* outs_w;
    int stride_words;
    int nmembers;
    int root;
    int vec;
    uint64_t lo, hi;
};

// Members are folded in batches of kFoldBatch: the batch's loads are all
// issued before its values are folded in rank order, so a lane keeps
// kFoldBatch * U vectors in flight instead of one.
constexpr int kFoldBatch = 4;

template <typename T, typename F>
__device__ inline void fold_slice(const FoldArgs& A, int t, int nt) {
    F f;
    const int P = A.nmembers;
    constexpr int V = sizeof(T) >= 16 ? 1 : (int)(16 / sizeof(T));
    constexpr int U = 4;
    constexpr int B = kFoldBatch;
    uint64_t lo = A.lo;
    const uint64_t hi = A.hi;
    const int r0 = A.root >= 0 ? A.root : 0;
    const int r1 = A.root >= 0 ? A.root + 1 : P;
    if (A.vec && V > 1) {
        struct alignas(16) Vt { T v[V]; };
        const uint64_t ngroups = (hi - lo) / V;
        const uint64_t gbase = lo / V;
        for (uint64_t g0 = (uint64_t)t; g0 < ngroups; g0 += (uint64_t)nt * U) {
            Vt acc[U];
            bool valid[U];
            {
                const Vt* p = (const Vt*)(uintptr_t)A.ins_w[0];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const uint64_t g = g0 + (uint64_t)u * nt;
                    valid[u] = g < ngroups;
                    if (valid[u]) acc[u] = p[gbase + g];
                }
            }
            for (int rb = 1; rb < P; rb += B) {
                const Vt* p[B];
#pragma unroll
                for (int k = 0; k < B; k++)
                    p[k] = rb + k < P ? (const Vt*)(uintptr_t)A.ins_w[(uint64_t)(rb + k) * A.stride_words] : nullptr;
                Vt x[B][U];
#pragma unroll
                for (int k = 0; k < B; k++)
#pragma unroll
                    for (int u = 0; u < U; u++)
                        if (p[k] && valid[u]) x[k][u] = p[k][gbase + g0 + (uint64_t)u * nt];
#pragma unroll
                for (int k = 0; k < B; k++)
                    if (p[k]) {
#pragma unroll
                        for (int u = 0; u < U; u++)
                            if (valid[u]) {
#pragma unroll
                                for (int e = 0; e < V; e++) acc[u].v[e] = f(x[k][u].v[e], acc[u].v[e]);
                            }
                    }
            }
            for (int r = r0; r < r1; r++) {
                Vt* o = (Vt*)(uintptr_t)A.outs_w[(uint64_t)r * A.stride_words];
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (valid[u]) o[gbase + g0 + (uint64_t)u * nt] = acc[u];
            }
        }
        lo += ngroups * V;
    }
    for (uint64_t e = lo + (uint64_t)t; e < hi; e += (uint64_t)nt) {
        T acc = ((const T*)(uintptr_t)A.ins_w[0])[e];
        for (int rb = 1; rb < P; rb += B) {
            T x[B];
#pragma unroll
            for (int k = 0; k < B; k++)
                if (rb + k < P) x[k] = ((const T*)(uintptr_t)A.ins_w[(uint64_t)(rb + k) * A.stride_words])[e];
#pragma unroll
            for (int k = 0; k < B; k++)
                if (rb + k < P) acc = f(x[k], acc);
        }
        for (int r = r0; r < r1; r++) ((T*)(uintptr_t)A.outs_w[(uint64_t)r * A.stride_words])[e] = acc;
    }
}

template <typename F>
__device__ inline void fold_unsigned_widths(int es, const FoldArgs& A, int t, int nt) {
    switch (es) {
    case 1: fold_slice<uint8_t, F>(A, t, nt); break;
    case 2: fold_slice<uint16_t, F>(A, t, nt); break;
    case 4: fold_slice<uint32_t, F>(A, t, nt); break;
    default: fold_slice<uint64_t, F>(A, t, nt); break;
    }
}

template <typename F>
__device__ inline void fold_ordered(int dt, const FoldArgs& A, int t, int nt) {
    const int es = type_size(dt);
    if (dt == MPI_FLOAT) return fold_slice<float, F>(A, t, nt);
    if (dt == MPI_DOUBLE || dt == MPI_LONG_DOUBLE) return fold_slice<double, F>(A, t, nt);
    if (type_signed_int(dt)) {
        switch (es) {
        case 1: return fold_slice<int8_t, F>(A, t, nt);
        case 2: return fold_slice<int16_t, F>(A, t, nt);
        case 4: return fold_slice<int32_t, F>(A, t, nt);
        default: return fold_slice<int64_t, F>(A, t, nt);
        }
    }
    fold_unsigned_widths<F>(es, A, t, nt);
}

template <typename F>
__device__ inline void fold_arith(int dt, const FoldArgs& A, int t, int nt) {
    if (dt == MPI_FLOAT) return fold_slice<float, F>(A, t, nt);
    if (dt == MPI_DOUBLE || dt == MPI_LONG_DOUBLE) return fold_slice<double, F>(A, t, nt);
    if (dt == MPI_C_COMPLEX || dt == MPI_C_FLOAT_COMPLEX) return fold_slice<cfloat2, F>(A, t, nt);
    if (dt == MPI_C_DOUBLE_COMPLEX || dt == MPI_C_LONG_DOUBLE_COMPLEX) return fold_slice<cdouble2, F>(A, t, nt);
    // Integer sum/product wrap around: unsigned arithmetic gives the same bits.
    fold_unsigned_widths<F>(type_size(dt), A, t, nt);
}

__device__ inline void fold_builtin(int dt, int op, const FoldArgs& A, int t, int nt) {
    switch (op) {
    case OP_MAX: fold_ordered<FMax>(dt, A, t, nt); break;
    case OP_MIN: fold_ordered<FMin>(dt, A, t, nt); break;
    case OP_SUM: fold_arith<FSum>(dt, A, t, nt); break;
    case OP_PROD: fold_arith<FProd>(dt, A, t, nt); break;
    case OP_LAND: fold_unsigned_widths<FLAnd>(type_size(dt), A, t, nt); break;
    case OP_LOR: fold_unsigned_widths<FLOr>(type_size(dt), A, t, nt); break;
    case OP_LXOR: fold_unsigned_widths<FLXor>(type_size(dt), A, t, nt); break;
    case OP_BAND: fold_unsigned_widths<FBAnd>(type_size(dt), A, t, nt); break;
    case OP_BOR: fold_unsigned_widths<FBOr>(type_size(dt), A, t, nt); break;
    case OP_BXOR: fold_unsigned_widths<FBXor>(type_size(dt), A, t, nt); break;
    default: break;
    }
}

}  // namespace rt
}  // namespace gpumpi

// User-defined operator (operators.cu:234-254): a device function pointer.
struct MPI_Op_impl {
    MPI_User_function* fn;
    int commute;
};
