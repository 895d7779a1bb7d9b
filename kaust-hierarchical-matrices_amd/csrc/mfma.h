// gfx950 matrix-core (MFMA) wrappers used by the multi-RHS H-matvec.
//
// All three shapes are 16x16 outputs; measured operand layouts on gfx950:
//   A (16 x K): lane l holds A[l % 16][KL*(l/16) + e], e < KL
//   B (K x 16): lane l holds B[KL*(l/16) + e][l % 16]
//   D (16 x 16): lane l, register v holds D[row(l/16, v)][l % 16] with
//                row = 4*(l/16) + v for f32 and f16 outputs, and
//                row = 4*v + (l/16) for the f64 16x16x4 output
// with KL = 1 (f64 16x16x4, f32 16x16x4), 4 (f16 16x16x16) or 8 (f16
// 16x16x32, Mfma32h: twice the products of the 16x16x16 shape per
// instruction on gfx950).
#pragma once
#include <hip/hip_runtime.h>

namespace hmx {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef _Float16 v4h __attribute__((ext_vector_type(4)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

template <class T>
struct Mfma;

template <>
struct Mfma<double> {
    using acc_t = double;
    using accv = v4d;
    using frag = double;
    static constexpr int KL = 1;
    __device__ static accv op(frag a, frag b, accv c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static frag make(const double* p) { return p[0]; }
    __device__ static int out_row(int lg, int v) { return 4 * v + lg; }
};

template <>
struct Mfma<float> {
    using acc_t = float;
    using accv = v4f;
    using frag = float;
    static constexpr int KL = 1;
    __device__ static accv op(frag a, frag b, accv c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static frag make(const float* p) { return p[0]; }
    __device__ static int out_row(int lg, int v) { return 4 * lg + v; }
};

template <>
struct Mfma<_Float16> {
    using acc_t = float;
    using accv = v4f;
    using frag = v4h;
    static constexpr int KL = 4;
    __device__ static accv op(frag a, frag b, accv c) { return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0); }
    __device__ static frag make(const _Float16* p) { return v4h{p[0], p[1], p[2], p[3]}; }
    __device__ static int out_row(int lg, int v) { return 4 * lg + v; }
};

// f16 16x16x32: a lane's 8 operand values per instruction
struct Mfma32h {
    using acc_t = float;
    using accv = v4f;
    using frag = v8h;
    static constexpr int KL = 8;
    __device__ static accv op(frag a, frag b, accv c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
    __device__ static frag make(const _Float16* p) { return v8h{p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]}; }
    __device__ static int out_row(int lg, int v) { return 4 * lg + v; }
};

}  // namespace hmx
