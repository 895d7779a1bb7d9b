// Device-side helpers shared by every gfx950 kernel of the library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hmx {

// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11).  Counter-based, so a
// value depends only on (counter, key): the CPU restatement draws the same bits.
constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

constexpr uint32_t kTagPoints = 0x504F0000u;
constexpr uint32_t kTagOmega = 0x4F4D0000u;
constexpr uint32_t kTagRhs = 0x52480000u;

__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)kPhiloxM0 * c[0];
        uint64_t p1 = (uint64_t)kPhiloxM1 * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
        k0 += kPhiloxW0; k1 += kPhiloxW1;
    }
}

// 53-bit uniform in [0, 1) from two 32-bit words
__host__ __device__ inline double u53(uint32_t a, uint32_t b) {
    return (double)(((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6)) * 0x1.0p-53;
}

// N(0,1) from one Philox block (Box-Muller, cosine branch).  (a, b, c) is the
// logical index: a = operation id (64 bit), b = sample column, c = row.
__device__ inline double gauss_at(uint64_t seed, uint32_t tag, uint64_t a, uint32_t b, uint32_t c) {
    uint32_t ctr[4] = {c, b, (uint32_t)a, tag ^ (uint32_t)(a >> 32)};
    philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
    double u1 = u53(ctr[0], ctr[1]) + 0x1.0p-53;
    double u2 = u53(ctr[2], ctr[3]);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// Exponential covariance K(x,y) = exp(-||x-y||/ell), same operation order as
// interaction() in the reference (src/TLRMatrixHelpers.cu:44-58).  Its
// diff += (x - y)*(x - y) is one fused multiply-add under nvcc's default
// contraction (--fmad=true); it is written as fma here, so every kernel of
// the library rounds it the same way whatever the compiler's contraction.
template <int D>
__device__ inline double kernel_eval(const double* __restrict__ a, const double* __restrict__ b, double ell) {
    double diff = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const double t = a[j] - b[j];
        diff = fma(t, t, diff);
    }
    return exp(-sqrt(diff) / ell);
}

// exp(x) for x <= 0 (the kernel's exponent -r/l), without the special-case
// handling a general exp needs.  Cody-Waite argument reduction: x = k ln2 + r
// with k = rint(x log2 e), |r| <= ln2/2, ln2 split into a head of 32
// significant bits (k * head is exact for |k| < 2^20) and a tail, so r carries
// no cancellation error; exp(r) by its Taylor series to degree 12 (remainder
// below 2e-16 relative on |r| <= ln2/2), in Horner form; 2^k by ldexp, which
// also flushes results below the double range to zero.  The split and the
// coefficients 1/n! follow from the definitions of ln2 and exp.
__device__ inline double exp_nonpos(double x) {
    constexpr double kLog2e = 1.4426950408889634;
    constexpr double kLn2Head = 0x1.62e42fee00000p-1;
    constexpr double kLn2Tail = 0x1.a39ef35793c76p-33;
    const double k = __builtin_rint(x * kLog2e);
    double r = fma(-k, kLn2Head, x);
    r = fma(-k, kLn2Tail, r);
    double p = 2.08767569878681e-09;  // 1/12!
    p = fma(p, r, 2.505210838544172e-08);
    p = fma(p, r, 2.755731922398589e-07);
    p = fma(p, r, 2.7557319223985893e-06);
    p = fma(p, r, 2.48015873015873e-05);
    p = fma(p, r, 0.0001984126984126984);
    p = fma(p, r, 0.001388888888888889);
    p = fma(p, r, 0.008333333333333333);
    p = fma(p, r, 0.041666666666666664);
    p = fma(p, r, 0.16666666666666666);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)k);
}

// sqrt(d) for d >= 0 within about 1 ulp: the hardware reciprocal square
// root (v_rsq_f64), one Goldschmidt step on g = d y and h = y/2 (each step
// squares the relative error) and one Newton correction g += (d - g^2) h.
// Without the rescaling a general sqrt needs for subnormal d (squared
// distances are 0 or far above the subnormal range): 10 instructions
// instead of 22.
__device__ inline double sqrt_nonneg(double d) {
    const double y = __builtin_amdgcn_rsq(d);
    double g = d * y, h = 0.5 * y;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    const double e = fma(-g, g, d);
    g = fma(e, h, g);
    return d > 0.0 ? g : 0.0;  // rsq(0) is infinite
}

// Same kernel with the division folded into a multiply by nil = -1/ell,
// sqrt_nonneg and exp_nonpos: for sampling loops whose values are not
// compared bit for bit.
template <int D>
__device__ inline double kernel_eval_scaled(const double* __restrict__ a, const double* __restrict__ b, double nil) {
    double diff = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const double t = a[j] - b[j];
        diff = fma(t, t, diff);
    }
    return exp_nonpos(nil * sqrt_nonneg(diff));
}

__device__ inline double kernel_eval_d(const double* a, const double* b, int d, double ell) {
    double diff = 0.0;
    for (int j = 0; j < d; ++j) {
        const double t = a[j] - b[j];
        diff = fma(t, t, diff);
    }
    return exp(-sqrt(diff) / ell);
}

// lane i of a 16-lane row receives lane (i ^ 1) / (i ^ 2) / (7 - i within 8) /
// (15 - i) of the same row (DPP quad_perm [1,0,3,2], [2,3,0,1],
// row_half_mirror, row_mirror)
template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)u, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ inline double lane_f64(double v, int lane) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)u, lane);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over the 64 lanes, the same value in every lane (wave-uniform): each
// DPP step adds two values that every lane of the pair adds in either order
// (a + b == b + a exactly), so every lane of a 16-lane row ends with the same
// row sum; the four row sums are then added in a fixed order.
__device__ inline double wave_sum_dpp(double v) {
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0x141>(v);
    v += dpp_f64<0x140>(v);
    return (lane_f64(v, 0) + lane_f64(v, 16)) + (lane_f64(v, 32) + lane_f64(v, 48));
}

// Morton helpers (src/helperKernels.cuh:87-125): column in even bits, row in odd bits
__host__ __device__ inline uint32_t spread16(uint32_t x) {
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}
__host__ __device__ inline uint32_t compact16(uint32_t x) {
    x &= 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0F0F0F0Fu;
    x = (x | (x >> 4)) & 0x00FF00FFu;
    x = (x | (x >> 8)) & 0x0000FFFFu;
    return x;
}
__host__ __device__ inline uint32_t morton_xy(uint32_t col, uint32_t row) { return spread16(col) | (spread16(row) << 1); }
__host__ __device__ inline uint32_t cm_to_mo(uint32_t S, uint32_t n) { return morton_xy(n / S, n % S); }

// Workgroup i of a launch starts on XCD i % 8 (8 XCDs, each with its own L2).
// Renumber so that every XCD walks one contiguous range of the work list:
// neighbouring leaves / tiles (which share right-hand-side blocks) then meet
// in the same L2.  A bijection on [0, n).
__device__ inline int xcd_local(int bid, int n) {
    const int q = n >> 3, r = n & 7, x = bid & 7, k = bid >> 3;
    return x * q + min(x, r) + k;
}

// Workgroup numbering in chunks: the hardware deals workgroup b to XCD b % 8;
// this remap hands XCD x the logical chunks x, x+8, x+16, ... of C
// consecutive workgroups, so neighbouring work (sharing right-hand-side
// rows) meets in one L2 while every XCD still draws from the whole range.
// A bijection on [0, 8*C*k); the grid is padded to such a size and logical
// numbers past the real count exit.
__device__ inline int xcd_chunked(int bid, int C) {
    const int x = bid & 7, slot = bid >> 3;
    return ((slot / C) * 8 + x) * C + slot % C;
}

// order-preserving map of a double onto uint64 (radix-sort key twiddle)
__device__ inline uint64_t dkey(double v) {
    uint64_t u = (uint64_t)__double_as_longlong(v);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

}  // namespace hmx
