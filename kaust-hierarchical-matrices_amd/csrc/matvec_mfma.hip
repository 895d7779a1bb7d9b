// Multi right-hand-side H-matvec on the matrix cores, FP64 (any number of
// right-hand sides, in passes of 16 or 64) and mixed FP32/FP16 (16 or 64).
//
//   phase 1  Z = V^T X per tile: one wave per (tile, 64 columns, 1024 rows),
//            D(16 cols x 16 rhs) += A(16 cols x K rows) B(K rows x 16 rhs);
//            the K index runs along V's rows.  FP64: V in place, one 32-byte
//            vector per operand and lane per four MFMAs; FP32/FP16: V copied
//            in fragment order and the right-hand sides in 16-byte row
//            blocks, so every load instruction reads whole cache lines, and
//            FP16 on the 16x16x32 shape.
//   reduce   chunk partials summed in a fixed order, scaled, converted to the
//            compute type and stored transposed ([rhs][column]).
//   phase 2  one wave per leaf (32 or 64 rows): Y_leaf = D X_leaf + sum_t U_t[leaf rows] Z_t,
//            A = the leaf's rows of U (fragment-ordered copy or the level's
//            column-major U in place), B = transposed Z.
//
// Mixed precision: the columns of every tile are split by a global energy
// budget.  A column j of tile t contributes ||u_j v_j^T||_F = ||u_j|| ||v_j||
// (= ||v_j|| for the ARA's orthonormal U, ||u_j|| for a mirrored tile's
// orthonormal V).  The columns whose squared norms sum to at most
// (target / (2*2^-11/sqrt(3)))^2 * ||H||_F^2 are stored in FP16 with
// per-column scales (max-abs of u_j and of v_j) and multiplied on the FP16
// matrix cores with FP32 accumulation, the rest (and the dense diagonal
// blocks) in FP32 on the FP32 matrix cores.  The right-hand sides go to FP16
// divided by their max-abs; FP16 z holds V^T x in those units times one
// plan-wide power of two (no z can overflow, whatever the scale of x), and
// phase 2 applies the right-hand-side scales to the FP32 sum of the FP16
// terms.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <type_traits>

#include "hmx_device.h"
#include "hmx_internal.h"
#include "mfma.h"

namespace hmx {

constexpr int kMfRows = 1024;  // rows per phase-1 item (FP64)

// rows per phase-1 item of the converted (FP32/FP16) groups: taller items
// write fewer partial sums of the tall tiles and the reduction reads fewer.
// Mixed 64 RHS at N = 2^20: 12.28 / 11.85 / 11.55 / 11.89 / 12.57 ms with
// 1024 / 2048 / 4096 / 8192 / 16384 rows (HMX_MIXED_ITEM_ROWS, a power of two
// >= 1024, for tuning)
static int mixed_item_rows(int64_t npad) {
    static const int v = [] {
        const char* e = std::getenv("HMX_MIXED_ITEM_ROWS");
        const int r = e ? std::atoi(e) : 0;
        return r >= kMfRows && (r & (r - 1)) == 0 ? r : 0;
    }();
    // a smaller plan keeps 1024-row items, so that there are enough of them
    // to fill the GPU (N = 65536: mixed 16 RHS 0.20 ms with 1024 rows, 0.37
    // with 4096)
    return v ? v : (npad >= ((int64_t)1 << 19) ? 4 * kMfRows : kMfRows);
}
constexpr int kMfCols = 64;    // columns per phase-1 item
constexpr int kPad = 16;       // rank padding of the row-major U copies
constexpr int kXSlices = 64;   // row slices of the right-hand-side max-abs scan

template <class T, int N>
__device__ inline void vload(const T* __restrict__ p, T (&out)[N]) {
    static_assert((N * sizeof(T)) % 16 == 0, "vector loads of whole 16-byte words");
    const uint4* q = reinterpret_cast<const uint4*>(p);
    uint4 tmp[N * sizeof(T) / 16];
#pragma unroll
    for (int i = 0; i < (int)(N * sizeof(T) / 16); ++i) tmp[i] = q[i];
    __builtin_memcpy(out, tmp, sizeof(tmp));
}

// the same for data read once (factors): non-temporal, so the streamed
// factors do not push the reused right-hand-side blocks and z out of L2
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
template <class T, int N>
__device__ inline void vload_nt(const T* __restrict__ p, T (&out)[N]) {
    static_assert((N * sizeof(T)) % 16 == 0, "vector loads of whole 16-byte words");
    const u32x4_t* q = reinterpret_cast<const u32x4_t*>(p);
    u32x4_t tmp[N * sizeof(T) / 16];
#pragma unroll
    for (int i = 0; i < (int)(N * sizeof(T) / 16); ++i) tmp[i] = __builtin_nontemporal_load(q + i);
    __builtin_memcpy(out, tmp, sizeof(tmp));
}

// factor loads: non-temporal when a lane's vector is 32 bytes, so the four
// lane groups of a row or column read whole 128-byte lines; a narrower
// vector leaves half a line that the next step reads, and a non-temporal
// load would let that line go and fetch it twice (measured on FP32 blocks)
template <class T, int N>
__device__ inline void vload_f(const T* __restrict__ p, T (&out)[N]) {
    if constexpr (N * sizeof(T) >= 32) vload_nt(p, out);
    else vload(p, out);
}

template <class T, int N>
__device__ inline void vzero(T (&out)[N]) {
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = (T)0;
}

// one 16-byte word into out[0 .. 16/sizeof(T)); the non-temporal form for
// data read once
template <class T>
__device__ inline void vload16(const T* __restrict__ p, T* out) {
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    __builtin_memcpy(out, &v, 16);
}
template <class T>
__device__ inline void vload16_nt(const T* __restrict__ p, T* out) {
    const u32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(p));
    __builtin_memcpy(out, &v, 16);
}

// z of a precision group: column c (padded, tiles start at multiples of 16)
// and right-hand side s.  Block c/VL of NR*VL elements holds the VL columns
// of a lane's B operand in phase 2, split into pieces of ZP columns: piece
// (c%VL)/ZP at piece*NR*ZP, right-hand side s at s*ZP.  FP32/FP16: 16-byte
// pieces, so a phase-2 load instruction (one piece per lane) reads 256
// contiguous bytes per lane group -- whole cache lines -- where one vector of
// VL columns per lane read half of each line per instruction (mixed phase 2
// 9.4 -> 7.2 ms with the U copies in fragment order, N = 2^20, 64
// right-hand sides).  FP64: one 32-byte piece (ZP = VL); the FP64 kernel
// runs one wave per SIMD and was slower with 16-byte pieces (37.9 -> 40.1 ms).
template <class T>
constexpr int z_piece() {
    return sizeof(T) == 8 ? 4 * Mfma<T>::KL : 16 / (int)sizeof(T);
}
template <class T, int NR>
__device__ inline size_t zt_index(int64_t c, int s) {
    constexpr int VL = 4 * Mfma<T>::KL, ZP = z_piece<T>();
    return (size_t)(c / VL) * (NR * VL) + (size_t)((c % VL) / ZP) * (NR * ZP) + (size_t)s * ZP + (size_t)(c % ZP);
}

// Element (r, c) of a run's row-major U copy (rows r < m, columns c < rpad,
// rpad a multiple of 16) in fragment order.  16-row blocks of 16*rpad
// elements, as in plain row-major order, so a leaf's rows start at r*rpad;
// inside a block, steps of SS = 16*KL columns (the last one w = rpad - c0
// wide), and inside a step the 16-byte pieces ordered (piece, lane group,
// row): lane (lg, lc)'s piece p of its VL columns c0 + VL*lg + [0, VL) of row
// lc sits at 16*c0 + p*16*E16*(w/VL) + (16*lg + lc)*E16.  Each load
// instruction of phase 2 then reads w*16 contiguous bytes.
template <class T>
__host__ __device__ inline size_t urm_index(int r, int c, int rpad) {
    constexpr int VL = 4 * Mfma<T>::KL, SS = 16 * Mfma<T>::KL, E16 = 16 / (int)sizeof(T);
    const int c0 = c / SS * SS, w = (rpad - c0 < SS ? rpad - c0 : SS), q = c - c0;
    return (size_t)(r / 16) * 16 * rpad + (size_t)16 * c0 + (size_t)((q % VL) / E16) * (16 * E16 * (w / VL)) +
           (size_t)(16 * (q / VL) + r % 16) * E16 + (size_t)(q % E16);
}

// Right-hand sides in blocks of 2^sh rows (the leaf size): block k holds the
// rows of leaf k for all NR right-hand sides, [s][row % 2^sh].  A leaf's (or
// a tile's) slice of X is then a few contiguous NR*2^sh-element blocks
// instead of NR strips n elements apart, one per right-hand side.
template <int NR>
__device__ inline size_t xb_index(int64_t row, int s, int sh) {
    return ((((size_t)(row >> sh)) * NR + (size_t)s) << sh) + (size_t)(row & ((1 << sh) - 1));
}

// ---- phase 1 ---------------------------------------------------------------
// SPLIT (T = FP16): the FP32 group of a mixed plan with V held as FP16 hi +
// lo (V/scale = hi + lo, lo at V + vLoOff) and the right-hand sides as FP16
// hi + lo in units of their max-abs (Xb, Xlo): three FP16 products per step
// (hi*hi + hi*lo + lo*hi, about 22 bits).  On the FP16 matrix cores this
// group reads the same bytes as an FP32 copy, where the FP32 16x16x4 shape
// at 64 right-hand sides is bound by the matrix cores.  TZ = FP32: z
// multiplied by the column and right-hand-side scales (true units); TZ =
// FP16: z times the column scale (which carries the group's power of two)
// stored as FP16 hi + lo, lo at Zt + zLoOff, in right-hand-side units.
template <class T, int CG, int NB, bool VF, bool SPLIT = false, class TZ = T>
__global__ void __launch_bounds__(256) k_mf_phase1(const MfItem* __restrict__ items, int nItems, int chunk,
                                                   const T* __restrict__ Xb, int sh,
                                                   typename Mfma<T>::acc_t* __restrict__ zpart,
                                                   TZ* __restrict__ Zt, const int32_t* __restrict__ zmap,
                                                   const float* __restrict__ colScaleV, int kvalid,
                                                   const T* __restrict__ Xlo, int64_t vLoOff,
                                                   const float* __restrict__ rhsScale, int64_t zLoOff) {
    static_assert(!SPLIT || std::is_same<T, _Float16>::value, "split operands: FP16 hi + lo");
    using M = Mfma<T>;
    // the products: FP16 on the 16x16x32 shape (a lane's 16 values of a
    // step feed two instructions instead of four; the k order inside a step
    // is the same for both operands), FP32/FP64 on M
    using MX = typename std::conditional<std::is_same<T, _Float16>::value, Mfma32h, M>::type;
    constexpr int KL = M::KL, VL = 4 * KL, NR = 16 * NB;
    const int lane = threadIdx.x & 63;
    // Items are sorted by the right-hand-side rows they read.  Chunks of
    // consecutive workgroups share an XCD (and its L2), the chunks spread
    // over all XCDs: items differ in cost (1024-row chunks of upper levels,
    // 64-row leaf tiles), so one contiguous range per XCD would not balance.
    const int wg = chunk > 0 ? xcd_chunked(blockIdx.x, chunk) : (int)blockIdx.x;
    // The item index is the same in every lane of the wave.  Declared so
    // (readfirstlane), the item loads into scalar registers and the loop runs
    // without exec masks: FP32/FP16 phase 1 8.3 -> 8.0 ms at N = 2^20, 64
    // right-hand sides; FP64 with 16 (256-cloud batch) 19.15 -> 18.95 ms per
    // call.  The FP64 kernel with 64 is scheduled worse that way (36.4 ->
    // 40.7 ms, measured) and keeps the per-lane index.
    const int wv = wg * 4 + (int)(threadIdx.x >> 6);
    const int wid = (sizeof(T) < 8 || NB == 1) ? __builtin_amdgcn_readfirstlane(wv) : wv;
    if (wid >= nItems) return;
    const int lg = lane >> 4, lc = lane & 15;
    using ACC = typename M::acc_t;
    const MfItem it = items[wid];
    const T* __restrict__ V = (const T*)it.v;
    typename M::accv acc[CG][NB];
#pragma unroll
    for (int g = 0; g < CG; ++g)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[g][nb] = typename M::accv{0, 0, 0, 0};
    // FP64 with 64 right-hand sides and FP16: software pipelined, the
    // operands of super-step r0 + 16*KL are in flight while the MFMAs of
    // super-step r0 run (FP64 36.4 ms vs 39.1 without, at N = 2^20; FP16
    // 4.4 -> 4.2 ms with the fragment-ordered copies).  FP64 with 16 and
    // FP32: no register double buffer -- the registers it frees buy more
    // waves per SIMD, which hide the load latency better (FP64 16 RHS 16.6 ->
    // 15.9 ms; FP32 slower with it).
    constexpr bool PF = !SPLIT && ((sizeof(T) == 8 && NB > 1) || sizeof(T) == 2);
    T a[CG][VL], bx[NB][VL], an[CG][VL], bn[NB][VL];
    constexpr int E16 = 16 / (int)sizeof(T), NP = VL / E16;
    auto load = [&](int r0, T (&A)[CG][VL], T (&B)[NB][VL]) {
        const int rr = r0 + VL * lg;
        const bool rok = rr < it.rows;
        if constexpr (sizeof(T) == 8 && VF) {
            // FP64 copies in fragment order: whole cache lines per load
            // instruction for V; the right-hand sides in blocks of a leaf
            const int pstride = 16 * E16 * (min(16 * KL, it.rows - r0) / VL);
#pragma unroll
            for (int g = 0; g < CG; ++g) {
                if (rok && 16 * g + lc < it.ncols) {
                    const T* vp = V + (size_t)g * 16 * it.ld + (size_t)16 * r0 + (16 * lg + lc) * E16;
#pragma unroll
                    for (int p = 0; p < NP; ++p) vload16_nt(vp + p * pstride, &A[g][p * E16]);
                } else {
                    vzero(A[g]);
                }
            }
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                if (rok && 16 * nb + lc < kvalid)
                    vload(Xb + xb_index<NR>((int64_t)it.xoff + rr, 16 * nb + lc, sh), B[nb]);
                else
                    vzero(B[nb]);
            }
        } else if constexpr (sizeof(T) < 8) {
            // FP32/FP16 copies in fragment order (V column c, row r at
            // urm_index(c, r, m) of its run; the item points at its first
            // column block and row chunk), the right-hand sides in blocks
            // of E16 rows (sh = log2(E16)): every load instruction reads
            // whole cache lines
            const int pstride = 16 * E16 * (min(16 * KL, it.rows - r0) / VL);
#pragma unroll
            for (int g = 0; g < CG; ++g) {
                if (rok && 16 * g + lc < it.ncols) {
                    const T* vp = V + (size_t)g * 16 * it.ld + (size_t)16 * r0 + (16 * lg + lc) * E16;
#pragma unroll
                    for (int p = 0; p < NP; ++p) vload16_nt(vp + p * pstride, &A[g][p * E16]);
                } else {
                    vzero(A[g]);
                }
            }
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                if (rok && 16 * nb + lc < kvalid) {
#pragma unroll
                    for (int p = 0; p < NP; ++p)
                        vload16(Xb + xb_index<NR>((int64_t)it.xoff + rr + p * E16, 16 * nb + lc, sh), &B[nb][p * E16]);
                } else {
                    vzero(B[nb]);
                }
            }
        } else {
#pragma unroll
            for (int g = 0; g < CG; ++g) {
                const int col = 16 * g + lc;
                if (rok && col < it.ncols) vload_f(V + (size_t)col * it.ld + rr, A[g]);
                else vzero(A[g]);
            }
            // right-hand sides past kvalid are zero in Xb: not loaded
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                if (rok && 16 * nb + lc < kvalid)
                    vload(Xb + xb_index<NR>((int64_t)it.xoff + rr, 16 * nb + lc, sh), B[nb]);
                else
                    vzero(B[nb]);
            }
        }
    };
    // SPLIT: the lo parts of V and of the right-hand sides, at the same
    // positions as the hi parts (into an/bn, which no prefetch uses then)
    auto load_lo = [&](int r0, T (&A)[CG][VL], T (&B)[NB][VL]) {
        const int rr = r0 + VL * lg;
        const bool rok = rr < it.rows;
        const int pstride = 16 * E16 * (min(16 * KL, it.rows - r0) / VL);
#pragma unroll
        for (int g = 0; g < CG; ++g) {
            if (rok && 16 * g + lc < it.ncols) {
                const T* vp = V + vLoOff + (size_t)g * 16 * it.ld + (size_t)16 * r0 + (16 * lg + lc) * E16;
#pragma unroll
                for (int p = 0; p < NP; ++p) vload16_nt(vp + p * pstride, &A[g][p * E16]);
            } else {
                vzero(A[g]);
            }
        }
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            if (rok && 16 * nb + lc < kvalid) {
#pragma unroll
                for (int p = 0; p < NP; ++p)
                    vload16(Xlo + xb_index<NR>((int64_t)it.xoff + rr + p * E16, 16 * nb + lc, sh), &B[nb][p * E16]);
            } else {
                vzero(B[nb]);
            }
        }
    };
    if (PF) load(0, a, bx);
    for (int r0 = 0; r0 < it.rows; r0 += 16 * KL) {
        const bool more = PF && r0 + 16 * KL < it.rows;
        if (!PF) load(r0, a, bx);
        if constexpr (SPLIT) load_lo(r0, an, bn);
        if (more) load(r0 + 16 * KL, an, bn);
#pragma unroll
        for (int e = 0; e < VL / MX::KL; ++e)
#pragma unroll
            for (int g = 0; g < CG; ++g)
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    acc[g][nb] = MX::op(MX::make(&a[g][e * MX::KL]), MX::make(&bx[nb][e * MX::KL]), acc[g][nb]);
        if constexpr (SPLIT) {
#pragma unroll
            for (int e = 0; e < VL / MX::KL; ++e)
#pragma unroll
                for (int g = 0; g < CG; ++g)
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        acc[g][nb] = MX::op(MX::make(&a[g][e * MX::KL]), MX::make(&bn[nb][e * MX::KL]), acc[g][nb]);
#pragma unroll
            for (int e = 0; e < VL / MX::KL; ++e)
#pragma unroll
                for (int g = 0; g < CG; ++g)
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        acc[g][nb] = MX::op(MX::make(&an[g][e * MX::KL]), MX::make(&bx[nb][e * MX::KL]), acc[g][nb]);
        }
        if (more) {
#pragma unroll
            for (int g = 0; g < CG; ++g)
#pragma unroll
                for (int q = 0; q < VL; ++q) a[g][q] = an[g][q];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                for (int q = 0; q < VL; ++q) bx[nb][q] = bn[nb][q];
        }
    }
    if constexpr (sizeof(T) < 8) {
        // FP32 / FP16 groups: lane (lg, lc) holds columns 16g + 4lg + v (v < 4)
        // for right-hand side 16nb + lc.  Their tiles' columns are padded to
        // a multiple of 4 in both orders, so the four belong to one tile and
        // sit at four consecutive z columns from a multiple of 4: one 8- or
        // 16-byte store per (g, nb), and a wave's stores fill whole 512-byte
        // stretches of z.
        if (it.direct) {
            typedef TZ tv4 __attribute__((ext_vector_type(4)));
#pragma unroll
            for (int g = 0; g < CG; ++g) {
                const int col0 = 16 * g + 4 * lg;
                if (col0 >= it.ncols) continue;
                const int64_t zc = zmap[it.zidx + col0];
                ACC cs[4] = {(ACC)1, (ACC)1, (ACC)1, (ACC)1};
                if (colScaleV) {
                    const float4 c4 = *reinterpret_cast<const float4*>(colScaleV + it.zidx + col0);
                    cs[0] = (ACC)c4.x; cs[1] = (ACC)c4.y; cs[2] = (ACC)c4.z; cs[3] = (ACC)c4.w;
                }
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) {
                    tv4 o;
                    if constexpr (SPLIT && sizeof(TZ) == 2) {
                        tv4 ol;
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const float zv = acc[g][nb][v] * cs[v];
                            o[v] = (TZ)zv;
                            ol[v] = (TZ)(zv - (float)o[v]);
                        }
                        *reinterpret_cast<tv4*>(Zt + zLoOff + zt_index<TZ, NR>(zc, 16 * nb + lc)) = ol;
                    } else if constexpr (SPLIT) {
                        const ACC rs = 16 * nb + lc < kvalid ? (ACC)rhsScale[16 * nb + lc] : (ACC)0;
#pragma unroll
                        for (int v = 0; v < 4; ++v) o[v] = (TZ)(acc[g][nb][v] * cs[v] * rs);
                    } else {
#pragma unroll
                        for (int v = 0; v < 4; ++v) o[v] = (TZ)(acc[g][nb][v] * cs[v]);
                    }
                    *reinterpret_cast<tv4*>(Zt + zt_index<TZ, NR>(zc, 16 * nb + lc)) = o;
                }
            }
            return;
        }
    }
    if (it.direct) {  // whole columns: scaled, converted, stored transposed ([rhs][column])
#pragma unroll
        for (int g = 0; g < CG; ++g)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int col = 16 * g + M::out_row(lg, v);
                if (col < it.ncols) {
                    // column scale by V order: no load waits on the z-column map
                    const int64_t zc = zmap[it.zidx + col];
                    const ACC cs = colScaleV ? (ACC)colScaleV[it.zidx + col] : (ACC)1;
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        Zt[zt_index<TZ, NR>(zc, 16 * nb + lc)] = (TZ)(acc[g][nb][v] * cs);
                }
            }
        return;
    }
#pragma unroll
    for (int g = 0; g < CG; ++g)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int col = 16 * g + M::out_row(lg, v);
            if (col < it.ncols)
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    zpart[(it.out + (int64_t)col * it.nch) * NR + 16 * nb + lc] = acc[g][nb][v];
        }
}

// ---- reduce partials, scale, convert, transpose ------------------------------
// nred columns split into row chunks: g-th such column is padded column
// redCol[g] with partials colPart[g] .. + colNpart[g].  One wave per column,
// lane = right-hand side (NR = 64; NR = 16 uses 16 lanes): the partials are
// summed by 8 interleaved accumulators folded in a fixed order, so the long
// columns of the top levels (up to m/1024 partials) are not one serial chain.
// SPL (T = FP16): z stored as FP16 hi + lo, lo at Zt + zLoOff
template <class T, class ACC, int NR, bool SPL = false>
__global__ void __launch_bounds__(256) k_mf_reduce(int64_t nred, int64_t ldz, const int64_t* __restrict__ redCol,
                                                   const int64_t* __restrict__ colPart,
                                                   const int32_t* __restrict__ colNpart,
                                                   const float* __restrict__ colScale, const ACC* __restrict__ zpart,
                                                   T* __restrict__ Zt, const float* __restrict__ rhsScale,
                                                   int64_t zLoOff) {
    const int64_t g = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int s = threadIdx.x & 63;
    if (g >= nred || s >= NR) return;
    const int np = colNpart[g];
    const ACC* p = zpart + colPart[g] * NR + s;
    ACC acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = 0;
    int q = 0;
    for (; q + 8 <= np; q += 8)
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[u] += p[(size_t)(q + u) * NR];
#pragma unroll
    for (int u = 0; u < 8; ++u)
        if (q + u < np) acc[u] += p[(size_t)(q + u) * NR];
    ACC v = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    const int64_t zc = redCol[g];
    if (colScale) v *= (ACC)colScale[zc];
    if (rhsScale) v *= (ACC)rhsScale[s];  // split FP32 group: right-hand sides in units of their max-abs
    const size_t zi = zt_index<T, NR>(zc, s);
    Zt[zi] = (T)v;
    if constexpr (SPL) Zt[(size_t)zLoOff + zi] = (T)(v - (ACC)(T)v);
}

// ---- phase 2 -----------------------------------------------------------------
// One wave per leaf of BL rows (RG = BL/16 row groups): all its rows, all
// right-hand sides, every contribution, accumulated in registers and stored
// once -- no partial sums across waves, no LDS.
//
// One wave's BL rows times a list of low-rank entries:
// acc[rg][nb] += U[16rg.., cols] Z[cols, rhs 16nb..].  The operands of the
// next 16*KL-column step (possibly of the next entry) load while the current
// step's MFMAs run.  rowOff: the wave's first row inside the entries' U.
template <int BL, bool UCM, class T, int NB>
__device__ inline void lr_rows(const MfEntry* __restrict__ ents, int e0, int e1, const T* __restrict__ Z, int rowOff,
                               int lg, int lc, int kvalid, typename Mfma<T>::accv (&acc)[BL / 16][NB]) {
    using M = Mfma<T>;
    using MX = typename std::conditional<std::is_same<T, _Float16>::value, Mfma32h, M>::type;  // as in phase 1
    constexpr int KL = M::KL, VL = 4 * KL, SS = 16 * KL, NR = 16 * NB, RG = BL / 16, E16 = 16 / (int)sizeof(T),
                  NP = VL / E16;
    if (e0 >= e1) return;
    T a[RG][VL], bz[NB][VL], an[RG][VL], bn[NB][VL];
    auto load = [&](const MfEntry& en, int c0, T (&A)[RG][VL], T (&B)[NB][VL]) {
        const int cc = c0 + VL * lg;
        if (cc < en.rpad) {
            // fragment-ordered copies (urm_index): piece p of lane (lg, lc)
            // at 16*c0 + p*16*E16*(w/VL) + (16lg + lc)*E16 of its 16-row block
            const int pstride = 16 * E16 * (min(SS, en.rpad - c0) / VL);
#pragma unroll
            for (int rg = 0; rg < RG; ++rg) {
                if (UCM) {  // column-major U in place: one element per column, zero past the rank
                    const T* U = (const T*)en.u + rowOff + 16 * rg + lc;
#pragma unroll
                    for (int q = 0; q < VL; ++q)
                        A[rg][q] = cc + q < en.rank ? __builtin_nontemporal_load(U + (size_t)(cc + q) * en.ld) : (T)0;
                } else {
                    const T* U = (const T*)en.u + (size_t)(rowOff / 16 + rg) * 16 * en.rpad + (size_t)16 * c0 +
                                 (16 * lg + lc) * E16;
#pragma unroll
                    for (int p = 0; p < NP; ++p) vload16_nt(U + p * pstride, &A[rg][p * E16]);
                }
            }
            // z of right-hand sides past kvalid is never used: not loaded
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                if (16 * nb + lc < kvalid) {
                    constexpr int ZP = z_piece<T>();
                    const T* zp = Z + zt_index<T, NR>(en.zcol + cc, 16 * nb + lc);
                    if constexpr (ZP == E16) {
#pragma unroll
                        for (int p = 0; p < VL / ZP; ++p) vload16(zp + p * NR * ZP, &B[nb][p * ZP]);
                    } else {
                        vload(zp, B[nb]);
                    }
                } else {
                    vzero(B[nb]);
                }
            }
        } else {
#pragma unroll
            for (int rg = 0; rg < RG; ++rg) vzero(A[rg]);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) vzero(B[nb]);
        }
    };
    int e = e0, c0 = 0;
    MfEntry en = ents[e];
    load(en, 0, a, bz);
    while (true) {
        int ne = e, nc = c0 + SS;
        MfEntry nen = en;
        if (nc >= en.rpad) {
            ne = e + 1;
            nc = 0;
            if (ne < e1) nen = ents[ne];
        }
        const bool more = ne < e1;
        if (more) load(nen, nc, an, bn);
#pragma unroll
        for (int q = 0; q < VL / MX::KL; ++q)
#pragma unroll
            for (int rg = 0; rg < RG; ++rg)
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    acc[rg][nb] = MX::op(MX::make(&a[rg][q * MX::KL]), MX::make(&bz[nb][q * MX::KL]), acc[rg][nb]);
        if (!more) break;
#pragma unroll
        for (int rg = 0; rg < RG; ++rg)
#pragma unroll
            for (int q = 0; q < VL; ++q) a[rg][q] = an[rg][q];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int q = 0; q < VL; ++q) bz[nb][q] = bn[nb][q];
        e = ne;
        c0 = nc;
        en = nen;
    }
}

// lr_rows for a split group: U and z each as FP16 hi + lo (the lo parts
// uLoOff / zLoOff elements past the hi parts, same fragment orders as an
// FP16 group), three FP16 16x16x32 products per step (hi*hi + hi*lo +
// lo*hi).  No register double buffer: the two operand sets take the
// registers a prefetch would.
template <int BL, int NB>
__device__ inline void lr_rows_split(const MfEntry* __restrict__ ents, int e0, int e1, const _Float16* __restrict__ Z,
                                     int64_t zLoOff, int64_t uLoOff, int rowOff, int lg, int lc, int kvalid,
                                     v4f (&acc)[BL / 16][NB]) {
    using M = Mfma<_Float16>;
    using MX = Mfma32h;
    constexpr int KL = M::KL, VL = 4 * KL, SS = 16 * KL, NR = 16 * NB, RG = BL / 16, E16 = 8, NP = VL / E16,
                  ZP = z_piece<_Float16>();
    static_assert(ZP == E16, "16-byte z pieces");
    _Float16 ah[RG][VL], al[RG][VL], bh[NB][VL], bl[NB][VL];
    for (int e = e0; e < e1; ++e) {
        const MfEntry en = ents[e];
        for (int c0 = 0; c0 < en.rpad; c0 += SS) {
            const int cc = c0 + VL * lg;
            if (cc < en.rpad) {
                const int pstride = 16 * E16 * (min(SS, en.rpad - c0) / VL);
#pragma unroll
                for (int rg = 0; rg < RG; ++rg) {
                    const _Float16* U = (const _Float16*)en.u + (size_t)(rowOff / 16 + rg) * 16 * en.rpad +
                                        (size_t)16 * c0 + (16 * lg + lc) * E16;
#pragma unroll
                    for (int p = 0; p < NP; ++p) {
                        vload16_nt(U + p * pstride, &ah[rg][p * E16]);
                        vload16_nt(U + uLoOff + p * pstride, &al[rg][p * E16]);
                    }
                }
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) {
                    if (16 * nb + lc < kvalid) {
                        const _Float16* zp = Z + zt_index<_Float16, NR>(en.zcol + cc, 16 * nb + lc);
#pragma unroll
                        for (int p = 0; p < VL / ZP; ++p) {
                            vload16(zp + p * NR * ZP, &bh[nb][p * ZP]);
                            vload16(zp + zLoOff + p * NR * ZP, &bl[nb][p * ZP]);
                        }
                    } else {
                        vzero(bh[nb]);
                        vzero(bl[nb]);
                    }
                }
            } else {
#pragma unroll
                for (int rg = 0; rg < RG; ++rg) {
                    vzero(ah[rg]);
                    vzero(al[rg]);
                }
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) {
                    vzero(bh[nb]);
                    vzero(bl[nb]);
                }
            }
#pragma unroll
            for (int q = 0; q < VL / MX::KL; ++q)
#pragma unroll
                for (int rg = 0; rg < RG; ++rg)
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        acc[rg][nb] = MX::op(MX::make(&ah[rg][q * MX::KL]), MX::make(&bh[nb][q * MX::KL]), acc[rg][nb]);
#pragma unroll
            for (int q = 0; q < VL / MX::KL; ++q)
#pragma unroll
                for (int rg = 0; rg < RG; ++rg)
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        acc[rg][nb] = MX::op(MX::make(&ah[rg][q * MX::KL]), MX::make(&bl[nb][q * MX::KL]), acc[rg][nb]);
#pragma unroll
            for (int q = 0; q < VL / MX::KL; ++q)
#pragma unroll
                for (int rg = 0; rg < RG; ++rg)
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        acc[rg][nb] = MX::op(MX::make(&al[rg][q * MX::KL]), MX::make(&bh[nb][q * MX::KL]), acc[rg][nb]);
        }
    }
}

// One wave's leaf times its dense BL x BL blocks (the diagonal block, then
// its near-field blocks), software pipelined over steps of 4*EPL columns.  A
// lane loads EPL = 32 bytes of a block row and of an X row per step, so the
// four lane groups of a row read one whole 128-byte line at once (an FP32 row
// read 64 bytes at a time was fetched twice).  Temporal loads: a block's
// lines are reused by the next step.
template <int BL, class TD, int NB>
__device__ inline void dense_rows(int leaf, const TD* __restrict__ Drm, const int32_t* __restrict__ nearPtr,
                                  const MfNear* __restrict__ nearB, const TD* __restrict__ Xb, int lg, int lc,
                                  int kvalid, typename Mfma<TD>::accv (&acc)[BL / 16][NB]) {
    using MD = Mfma<TD>;
    static_assert(MD::KL == 1, "dense blocks on the 16x16x4 shapes");
    constexpr int NR = 16 * NB, EPL = 32 / sizeof(TD), SPB = BL / (4 * EPL), RG = BL / 16;
    static_assert(SPB >= 1, "a step spans at most one block row");
    const int n0 = nearPtr ? nearPtr[leaf] : 0, n1 = nearPtr ? nearPtr[leaf + 1] : 0;
    const int total = (n1 - n0 + 1) * SPB;
    TD a[RG][EPL], bx[NB][EPL], an[RG][EPL], bn[NB][EPL];
    auto dload = [&](int idx, TD (&A)[RG][EPL], TD (&B)[NB][EPL]) {
        const int q = n0 - 1 + idx / SPB;
        const int cc = (idx % SPB) * 4 * EPL + EPL * lg;
        const TD* Dq = q < n0 ? Drm + (size_t)leaf * BL * BL : (const TD*)nearB[q].d;
        const int cl = q < n0 ? leaf : nearB[q].colLeaf;
#pragma unroll
        for (int rg = 0; rg < RG; ++rg) vload(Dq + (16 * rg + lc) * BL + cc, A[rg]);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {  // right-hand sides past kvalid are zero in Xb: not loaded
            if (16 * nb + lc < kvalid) {
                if constexpr (sizeof(TD) < 8) {  // FP32 right-hand sides in blocks of 4 rows
#pragma unroll
                    for (int p = 0; p < EPL / 4; ++p)
                        vload16(Xb + xb_index<NR>((int64_t)cl * BL + cc + 4 * p, 16 * nb + lc, 2), &B[nb][4 * p]);
                } else {  // FP64: blocks of one leaf
                    vload(Xb + ((size_t)cl * NR + 16 * nb + lc) * BL + cc, B[nb]);
                }
            } else {
                vzero(B[nb]);
            }
        }
    };
    dload(0, a, bx);
    for (int idx = 0; idx < total; ++idx) {
        const bool more = idx + 1 < total;
        if (more) dload(idx + 1, an, bn);
#pragma unroll
        for (int e = 0; e < EPL; ++e)
#pragma unroll
            for (int rg = 0; rg < RG; ++rg)
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) acc[rg][nb] = MD::op(a[rg][e], bx[nb][e], acc[rg][nb]);
        if (more) {
#pragma unroll
            for (int rg = 0; rg < RG; ++rg)
#pragma unroll
                for (int e = 0; e < EPL; ++e) a[rg][e] = an[rg][e];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                for (int e = 0; e < EPL; ++e) bx[nb][e] = bn[nb][e];
        }
    }
}

// Mixed plans with the points at hand: one wave's leaf times the kernel
// blocks K(x_leaf, x_j) of its diagonal (j = leaf) and near-field column
// leaves j, evaluated on the fly instead of read.  Points are held in FP32
// relative to their leaf's first point (PL), with those origins in FP64 (O):
// x_r - x_c = (x_r - o_i) + (o_i - o_j) - (x_c - o_j), where o_i - o_j is
// formed in FP64 once per block and every term is small, so the differences
// carry errors near 1e-8 absolute; sqrt and exp in FP32 on the hardware
// instructions (exp(-r/l) = 2^(r * nlog2e)).  K is split into
// FP16 K_hi + K_lo and the right-hand sides (in units of their max-abs) into
// X_hi + X_lo, so three FP16 16x16x32 products per step (K_hi X_hi + K_hi X_lo
// + K_lo X_hi) carry about 22 bits; the K_lo X_lo term is below 2^-22 of the
// product.  Lane (lg, lc) evaluates K at rows 16rg + lc and columns
// 32ks + 8lg + e, e < 8: the A operand layout of the 16x16x32 shape; its B
// operand is one 16-byte block of 8 right-hand-side rows.  Padding
// positions hold the point (0, 0, 0): finite kernel values that meet zero
// right-hand-side entries or land in discarded rows.
template <int BL, int NB>
__device__ inline void near_otf(int leaf, int w, const float4* __restrict__ PL, const double4* __restrict__ O,
                                float nlog2e, const int32_t* __restrict__ gnearPtr, const MfNear* __restrict__ gnear,
                                const _Float16* __restrict__ Xh, const _Float16* __restrict__ Xl, int lg, int lc,
                                int kvalid, v4f (&acc)[BL / 16][NB]) {
    constexpr int RG = BL / 16, NR = 16 * NB, SH = 3;  // right-hand sides in blocks of 8 rows
    float xl[RG][3];
#pragma unroll
    for (int rg = 0; rg < RG; ++rg) {
        const float4 p = PL[(size_t)leaf * BL + 16 * rg + lc];
        xl[rg][0] = p.x; xl[rg][1] = p.y; xl[rg][2] = p.z;
    }
    const double4 oi = O[leaf];
    const int grp = leaf >> 2;
    const int n0 = gnearPtr[grp], n1 = gnearPtr[grp + 1];
    for (int q = n0; q < n1; ++q) {
        const MfNear e = gnear[q];
        if (!((__builtin_amdgcn_readfirstlane(e.pad) >> w) & 1)) continue;  // not this wave's block
        const int cl = __builtin_amdgcn_readfirstlane(e.colLeaf);
        const double4 oj = O[cl];
        const float d0 = (float)(oi.x - oj.x), d1 = (float)(oi.y - oj.y), d2 = (float)(oi.z - oj.z);
        float xr[RG][3];
#pragma unroll
        for (int rg = 0; rg < RG; ++rg) {
            xr[rg][0] = xl[rg][0] + d0; xr[rg][1] = xl[rg][1] + d1; xr[rg][2] = xl[rg][2] + d2;
        }
        for (int ks = 0; ks < BL / 32; ++ks) {
            const int c0 = cl * BL + 32 * ks + 8 * lg;  // this lane's 8 columns (padded positions)
            float xc[8][3];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float4 p = PL[c0 + e];
                xc[e][0] = p.x; xc[e][1] = p.y; xc[e][2] = p.z;
            }
            v8h bh[NB], bl[NB];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                if (16 * nb + lc < kvalid) {
                    const size_t o = xb_index<NR>((int64_t)c0, 16 * nb + lc, SH);
                    bh[nb] = *reinterpret_cast<const v8h*>(Xh + o);
                    bl[nb] = *reinterpret_cast<const v8h*>(Xl + o);
                } else {
                    bh[nb] = v8h{0, 0, 0, 0, 0, 0, 0, 0};
                    bl[nb] = v8h{0, 0, 0, 0, 0, 0, 0, 0};
                }
            }
#pragma unroll
            for (int rg = 0; rg < RG; ++rg) {
                v8h ah, al;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float dx = xr[rg][0] - xc[e][0], dy = xr[rg][1] - xc[e][1], dz = xr[rg][2] - xc[e][2];
                    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    // the hardware square root and 2^x (v_sqrt_f32, v_exp_f32:
                    // within 1 ulp, results below 2^-126 flushed to zero), not
                    // the correctly rounded library sequences: 4 instructions
                    // instead of 27, and the product carries 22 bits anyway
                    const float k = __builtin_amdgcn_exp2f(nlog2e * __builtin_amdgcn_sqrtf(r2));
                    const _Float16 hi = (_Float16)k;
                    ah[e] = hi;
                    al[e] = (_Float16)(k - (float)hi);
                }
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) {
                    acc[rg][nb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[nb], acc[rg][nb], 0, 0, 0);
                    acc[rg][nb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[nb], acc[rg][nb], 0, 0, 0);
                    acc[rg][nb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[nb], acc[rg][nb], 0, 0, 0);
                }
            }
        }
    }
}

// One workgroup of 4 waves per group of 4 consecutive leaves (S leaves in
// all); wave w owns leaf 4g+w: its dense blocks and its entries (ptr, per
// leaf), then -- grouped plans (64-row leaves) -- the tiles spanning the
// whole group (gptr, per group; null otherwise).  The four waves walk the
// shared list in the same order, so a tile's z block comes from HBM once per
// group and reaches the other waves from the caches.  Mixed plans: the FP16
// group (B) goes first into the cleared accumulators, which are then
// multiplied by loMul (2^-E, the plan's z exponent); with the points at hand
// (P) the diagonal and near-field blocks are evaluated on the fly and added in
// the same units (right-hand sides divided by their max-abs); then the sum is
// multiplied by loScale[rhs] (those max-abs values) before the stored dense
// blocks (without P) and the FP32 group add theirs.  SPA: the FP32 group
// held as FP16 hi + lo (U, and z in right-hand-side units times 2^Ea) goes
// first instead, its sum multiplied by hiMul = 2^(E - Ea) into the FP16
// group's units.
template <int BL, bool UCM, class TD, class TA, class TB, int NB, class TO, bool SPA = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SPA ? 2 : 1))) k_mf_phase2g(int S, const int32_t* __restrict__ ptrA, const MfEntry* __restrict__ entA,
                                                    const int32_t* __restrict__ gptrA,
                                                    const MfEntry* __restrict__ gentA, const TA* __restrict__ ZA,
                                                    const int32_t* __restrict__ ptrB, const MfEntry* __restrict__ entB,
                                                    const int32_t* __restrict__ gptrB,
                                                    const MfEntry* __restrict__ gentB, const TB* __restrict__ ZB,
                                                    const TD* __restrict__ Drm, const int32_t* __restrict__ nearPtr,
                                                    const MfNear* __restrict__ nearB, const TD* __restrict__ Xb,
                                                    TO* __restrict__ Y, int ldy, int kvalid,
                                                    const float* __restrict__ loScale, float loMul,
                                                    const float4* __restrict__ P, const double4* __restrict__ O,
                                                    float nlog2e,
                                                    const _Float16* __restrict__ Xh,
                                                    const _Float16* __restrict__ Xl, int64_t uLoA, int64_t zLoA,
                                                    float hiMul) {
    static_assert(!SPA || std::is_same<TA, _Float16>::value, "split group: FP16 hi + lo");
    using MA = Mfma<TA>;
    using accv = typename MA::accv;
    static_assert(sizeof(typename Mfma<TB>::acc_t) == sizeof(typename MA::acc_t) &&
                      sizeof(typename Mfma<TD>::acc_t) == sizeof(typename MA::acc_t),
                  "one accumulator type per launch");
    constexpr int RG = BL / 16;
    const int grp = xcd_local(blockIdx.x, gridDim.x);
    // wave-uniform leaf (readfirstlane): scalar entry loads, no exec masks on
    // the entry loops.  FP64 phase 2 39.2 -> 38.1 ms at N = 2^20, 64
    // right-hand sides; the mixed kernel is slower that way (9.7 -> 10.2 ms,
    // measured) and keeps the per-lane index.
    const int w = sizeof(TA) == 8 ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63, lg = lane >> 4, lc = lane & 15;
    const int leaf = 4 * grp + w;
    if (leaf >= S) return;  // last group of a plan whose leaf count is not a multiple of 4 (no barriers below)
    accv acc[RG][NB];
#pragma unroll
    for (int rg = 0; rg < RG; ++rg)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[rg][nb] = accv{0, 0, 0, 0};
    if constexpr (SPA) {
        lr_rows_split<BL, NB>(entA, ptrA[leaf], ptrA[leaf + 1], ZA, zLoA, uLoA, 0, lg, lc, kvalid, acc);
        if (gptrA)
            lr_rows_split<BL, NB>(gentA, gptrA[grp], gptrA[grp + 1], ZA, zLoA, uLoA, BL * w, lg, lc, kvalid, acc);
#pragma unroll
        for (int rg = 0; rg < RG; ++rg)
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) acc[rg][nb] *= hiMul;
    }
    if (ptrB) {
        lr_rows<BL, false, TB, NB>(entB, ptrB[leaf], ptrB[leaf + 1], ZB, 0, lg, lc, kvalid, acc);
        if (gptrB) lr_rows<BL, false, TB, NB>(gentB, gptrB[grp], gptrB[grp + 1], ZB, BL * w, lg, lc, kvalid, acc);
#pragma unroll
        for (int rg = 0; rg < RG; ++rg)
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) acc[rg][nb] *= (typename MA::acc_t)loMul;
    }
    if constexpr (std::is_same<TB, _Float16>::value && std::is_same<typename MA::accv, v4f>::value) {
        if (P) near_otf<BL, NB>(leaf, w, P, O, nlog2e, nearPtr, nearB, Xh, Xl, lg, lc, kvalid, acc);
    }
    if (ptrB) {
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const typename MA::acc_t f =
                16 * nb + lc < kvalid ? (typename MA::acc_t)loScale[16 * nb + lc] : (typename MA::acc_t)0;
#pragma unroll
            for (int rg = 0; rg < RG; ++rg) acc[rg][nb] *= f;
        }
    }
    if (!P) dense_rows<BL, TD, NB>(leaf, Drm, nearPtr, nearB, Xb, lg, lc, kvalid, acc);
    if constexpr (!SPA) {
        lr_rows<BL, UCM, TA, NB>(entA, ptrA[leaf], ptrA[leaf + 1], ZA, 0, lg, lc, kvalid, acc);
        if (gptrA) lr_rows<BL, UCM, TA, NB>(gentA, gptrA[grp], gptrA[grp + 1], ZA, BL * w, lg, lc, kvalid, acc);
    }
#pragma unroll
    for (int rg = 0; rg < RG; ++rg)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                if (16 * nb + lc < kvalid)
                    Y[(size_t)(16 * nb + lc) * ldy + (size_t)leaf * BL + 16 * rg + MA::out_row(lg, v)] = (TO)acc[rg][nb][v];
}

// ---- right-hand-side conversion (mixed) ---------------------------------------
// max |X[s][i]| over slice p of the rows: part[s*P + p], grid (P, right-hand sides)
__global__ void __launch_bounds__(256) k_colmaxabs_part(int n, const float* __restrict__ X, int64_t ldx,
                                                        float* __restrict__ part) {
    __shared__ float red[256];
    const int P = gridDim.x;
    const int64_t per = ((int64_t)n + P - 1) / P;
    const int64_t i0 = (int64_t)blockIdx.x * per, i1 = min((int64_t)n, i0 + per);
    const float* x = X + (size_t)blockIdx.y * ldx;
    float m = 0.f;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) m = fmaxf(m, fabsf(x[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < (unsigned)w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.y * P + blockIdx.x] = red[0];
}

// sx[s] = max over the P slices (1 for an all-zero right-hand side); one wave per s
__global__ void __launch_bounds__(64) k_colmaxabs_fin(int P, const float* __restrict__ part, float* __restrict__ sx) {
    float m = 0.f;
    for (int p = threadIdx.x; p < P; p += 64) m = fmaxf(m, part[(size_t)blockIdx.x * P + p]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if (threadIdx.x == 0) sx[blockIdx.x] = m > 0.f ? m : 1.f;
}

// X (right-hand side s at X + s*ldx) into the blocked layout of the padded
// leaf space (npad positions): position p takes entry gmap[p] (p itself
// without a map; zero on padding, gmap[p] < 0), right-hand sides s >= kvalid
// are zero; divided by the per-right-hand-side scale sx when one is given
template <class TI, class TO>
__global__ void __launch_bounds__(256) k_x_blocked(int64_t total, int64_t npad, int NR, int sh,
                                                   const TI* __restrict__ X, int64_t ldx, int kvalid,
                                                   const int32_t* __restrict__ gmap, const float* __restrict__ sx,
                                                   TO* __restrict__ Xb) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= total) return;
    const int s = (int)(e / npad);
    const int64_t p = e % npad;
    const size_t o = ((((size_t)(p >> sh)) * NR + (size_t)s) << sh) + (size_t)(p & ((1 << sh) - 1));
    TO v = (TO)0;
    if (s < kvalid) {
        const int64_t g = gmap ? (int64_t)gmap[p] : p;
        if (g >= 0) v = sx ? (TO)(X[s * ldx + g] / sx[s]) : (TO)X[s * ldx + g];
    }
    Xb[o] = v;
}

// Points of the padded leaf space (x, y, z, w as doubles) relative to their
// leaf's first position: PL[p] = (x - o, y - o, z - o) in FP32, O[leaf] = o in FP64
__global__ void __launch_bounds__(256) k_local_points(int64_t npad, int b, const double4* __restrict__ P,
                                                      float4* __restrict__ PL, double4* __restrict__ O) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= npad) return;
    const int64_t leaf = e / b;
    const double4 o = P[leaf * b];
    const double4 p = P[e];
    PL[e] = make_float4((float)(p.x - o.x), (float)(p.y - o.y), (float)(p.z - o.z), 0.f);
    if (e % b == 0) O[leaf] = o;
}

// All right-hand-side copies of a mixed plan from one read of X: FP32 in
// blocks of 4 rows (Xb), FP16 hi = fp16(x/sx) and, when Xl is given, lo =
// fp16(x/sx - hi) in blocks of 8 rows (zero on padding and past kvalid).
// One workgroup per 32 padded positions: X is read coalesced along the rows
// into LDS, and each output's 32*NR-element stretch of the blocked layout
// (contiguous for a 32-row chunk) is written in order with 16-byte stores.
__global__ void __launch_bounds__(256) k_x_mixed(int64_t npad, int NR, const float* __restrict__ X, int64_t ldx,
                                                 int kvalid, const int32_t* __restrict__ gmap,
                                                 const float* __restrict__ sx, float* __restrict__ Xb,
                                                 _Float16* __restrict__ Xh, _Float16* __restrict__ Xl) {
    __shared__ float tile[64][33];
    const int64_t p0 = (int64_t)blockIdx.x * 32;
    const int t = threadIdx.x;
    const int pp = t & 31;
    const int64_t g = gmap ? (int64_t)gmap[p0 + pp] : p0 + pp;
    for (int s = t >> 5; s < NR; s += 8) tile[s][pp] = (s < kvalid && g >= 0) ? X[s * ldx + g] : 0.f;
    __syncthreads();
    const size_t base = (size_t)p0 * NR;
    // FP32: element idx of the stretch = block (idx/4)/NR of 4 rows, right-hand side (idx/4)%NR
    for (int q = t; q < 8 * NR; q += 256) {
        const int s = q % NR, r0 = 4 * (q / NR);
        *reinterpret_cast<float4*>(Xb + base + 4 * (size_t)q) =
            make_float4(tile[s][r0], tile[s][r0 + 1], tile[s][r0 + 2], tile[s][r0 + 3]);
    }
    // FP16: blocks of 8 rows
    for (int q = t; q < 4 * NR; q += 256) {
        const int s = q % NR, r0 = 8 * (q / NR);
        const float inv = s < kvalid ? sx[s] : 1.f;
        v8h hi, lo;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float v = tile[s][r0 + j] / inv;
            hi[j] = (_Float16)v;
            lo[j] = (_Float16)(v - (float)hi[j]);
        }
        *reinterpret_cast<v8h*>(Xh + base + 8 * (size_t)q) = hi;
        if (Xl) *reinterpret_cast<v8h*>(Xl + base + 8 * (size_t)q) = lo;
    }
}

// product in the padded space (Yp, ld npad) back to the caller's entries
template <class T>
__global__ void __launch_bounds__(256) k_y_scatter(int64_t npad, int k, const int32_t* __restrict__ gmap,
                                                   const T* __restrict__ Yp, T* __restrict__ Y, int64_t ldy) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= npad * k) return;
    const int64_t s = e / npad, p = e % npad;
    const int g = gmap[p];
    if (g >= 0) Y[s * ldy + g] = Yp[e];
}

// ---- plan construction kernels -----------------------------------------------
struct ColDesc {
    const double* u;
    const double* v;
    int32_t m, pad;
};

// per column: ||u_j v_j^T||_F^2 = ||u||^2 ||v||^2 (||v||^2 for the ARA's
// orthonormal U, ||u||^2 for a mirrored tile's orthonormal V), max|u|,
// max|v|  (one wave per column)
__global__ void __launch_bounds__(256) k_colstats(int64_t ncols, const ColDesc* __restrict__ cols,
                                                  double* __restrict__ e2, double* __restrict__ umax,
                                                  double* __restrict__ vmax) {
    const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= ncols) return;
    const int lane = threadIdx.x & 63;
    const ColDesc d = cols[c];
    double su = 0, sv = 0, mu = 0, mv = 0;
    for (int i = lane; i < d.m; i += 64) {
        const double u = d.u[i], v = d.v[i];
        su += u * u;
        sv += v * v;
        mu = fmax(mu, fabs(u));
        mv = fmax(mv, fabs(v));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        su += __shfl_xor(su, off, 64);
        sv += __shfl_xor(sv, off, 64);
        mu = fmax(mu, __shfl_xor(mu, off, 64));
        mv = fmax(mv, __shfl_xor(mv, off, 64));
    }
    if (lane == 0) { e2[c] = su * sv; umax[c] = mu; vmax[c] = mv; }
}

// V column gather + scale + convert: dst column c = src / scale (one wave per column)
// dst is the run's first element and the column sits at urm_index(col, i, m)
// (fragment order, rows and columns of U's order exchanged)
struct CvCol {
    const double* src;
    int64_t dst;   // element offset in the group's V buffer
    int32_t m;
    float scale;
    int32_t col;   // column inside its run (fragment-ordered copies)
    int32_t pad;
};
template <class T>
__global__ void __launch_bounds__(256) k_cvt_cols(int64_t n, const CvCol* __restrict__ cols, T* __restrict__ V) {
    const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= n) return;
    const int lane = threadIdx.x & 63;
    const CvCol d = cols[c];
    const double inv = 1.0 / (double)d.scale;
    for (int i = lane; i < d.m; i += 64) {
        const T v = d.src ? (T)(d.src[i] * inv) : (T)0;
        V[d.dst + urm_index<T>(d.col, i, d.m)] = v;
    }
}

// The split copy of an FP32 group's V: hi = fp16(v/scale), lo = fp16(v/scale
// - hi) (formed in FP64), both in the FP16 fragment order, lo at loOff
__global__ void __launch_bounds__(256) k_cvt_cols_split(int64_t n, const CvCol* __restrict__ cols,
                                                        _Float16* __restrict__ V, int64_t loOff) {
    const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= n) return;
    const int lane = threadIdx.x & 63;
    const CvCol d = cols[c];
    const double inv = 1.0 / (double)d.scale;
    for (int i = lane; i < d.m; i += 64) {
        const double x = d.src ? d.src[i] * inv : 0.0;
        const _Float16 hi = (_Float16)x;
        const size_t o = (size_t)d.dst + urm_index<_Float16>(d.col, i, d.m);
        V[o] = hi;
        V[(size_t)loOff + o] = (_Float16)(x - (double)hi);
    }
}

// row-major U copy with zero padding, in fragment order (urm_index): one
// block per (tile, 32-row tile, 32-col tile)
struct TrTile {
    int64_t colDesc;  // first CvCol of the tile (column c of the tile = colDesc + c)
    int64_t dst;      // element offset of the tile's row-major copy
    int32_t m, cnt, rpad, r0, c0, pad;
};
template <class T>
__global__ void __launch_bounds__(256) k_cvt_rowmajor(int64_t n, const TrTile* __restrict__ tiles,
                                                      const CvCol* __restrict__ cols, T* __restrict__ U) {
    __shared__ double sh[32][33];
    const TrTile t = tiles[blockIdx.x];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int cc = ty; cc < 32; cc += 8) {  // read 32 columns (contiguous rows)
        const int c = t.c0 + cc, r = t.r0 + tx;
        double v = 0.0;
        if (c < t.cnt && r < t.m) {
            const CvCol d = cols[t.colDesc + c];
            v = d.src ? d.src[r] / (double)d.scale : 0.0;
        }
        sh[cc][tx] = v;
    }
    __syncthreads();
    for (int rr = ty; rr < 32; rr += 8) {  // write 32 rows (contiguous columns)
        const int r = t.r0 + rr, c = t.c0 + tx;
        if (r < t.m && c < t.rpad) U[t.dst + urm_index<T>(r, c, t.rpad)] = (T)sh[tx][rr];
    }
}

// the same for a split group's U: FP16 hi + lo (formed in FP64), lo at loOff
__global__ void __launch_bounds__(256) k_cvt_rowmajor_split(int64_t n, const TrTile* __restrict__ tiles,
                                                            const CvCol* __restrict__ cols, _Float16* __restrict__ U,
                                                            int64_t loOff) {
    __shared__ double sh[32][33];
    const TrTile t = tiles[blockIdx.x];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int cc = ty; cc < 32; cc += 8) {
        const int c = t.c0 + cc, r = t.r0 + tx;
        double v = 0.0;
        if (c < t.cnt && r < t.m) {
            const CvCol d = cols[t.colDesc + c];
            v = d.src ? d.src[r] / (double)d.scale : 0.0;
        }
        sh[cc][tx] = v;
    }
    __syncthreads();
    for (int rr = ty; rr < 32; rr += 8) {
        const int r = t.r0 + rr, c = t.c0 + tx;
        if (r < t.m && c < t.rpad) {
            const double x = sh[tx][rr];
            const _Float16 hi = (_Float16)x;
            const size_t o = (size_t)t.dst + urm_index<_Float16>(r, c, t.rpad);
            U[o] = hi;
            U[(size_t)loOff + o] = (_Float16)(x - (double)hi);
        }
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_diag_rowmajor(int64_t total, int b, const double* __restrict__ D,
                                                       T* __restrict__ Drm) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= total) return;
    const int64_t bb = (int64_t)b * b;
    const int64_t leaf = e / bb;
    const int r = (int)((e % bb) / b), c = (int)(e % b);
    Drm[e] = (T)D[leaf * bb + (int64_t)c * b + r];  // D is column-major per leaf
}

// dense leaf-level blocks: out[dst + row*b + col] = sum_a U[a*b + row] V[a*b + col]
struct NearSrc {
    const double* U;
    const double* V;
    int64_t dst;
    int32_t r, pad;
};
template <class T>
__global__ void __launch_bounds__(256) k_near_dense(int b, const NearSrc* __restrict__ src, T* __restrict__ out) {
    const NearSrc t = src[blockIdx.x];
    for (int e = threadIdx.x; e < b * b; e += 256) {
        const int row = e / b, col = e % b;
        double v = 0.0;
        for (int a = 0; a < t.r; ++a) v = fma(t.U[(size_t)a * b + row], t.V[(size_t)a * b + col], v);
        out[t.dst + e] = (T)v;
    }
}

// ---- host side ------------------------------------------------------------------
namespace {

struct TileRef {
    const HLevel* lv;
    int level, t, r, m;
    uint32_t inCl, outCl;
    int64_t leafBase;  // first leaf of the tile's H-matrix in the plan's leaf space
};

std::vector<TileRef> all_tiles(const HMatrix& h, int64_t leafBase = 0) {
    std::vector<TileRef> v;
    for (int l = 1; l < h.numLevels; ++l) {
        const HLevel& lv = h.levels[l - 1];
        const auto& tiles = h.structure->tileIndices[l - 1];
        for (int t = 0; t < lv.numTiles; ++t) {
            if (lv.h_ranks[t] == 0) continue;
            v.push_back(TileRef{&lv, l, t, lv.h_ranks[t], lv.m, compact16((uint32_t)tiles[t]),
                                compact16((uint32_t)tiles[t] >> 1), leafBase});
        }
    }
    return v;
}

template <class T>
void upload(Context& c, DBuf<T>& d, const std::vector<T>& h) {
    d.alloc(std::max<size_t>(1, h.size()));
    if (!h.empty()) HMX_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, c.stream));
}

size_t tsize(int type) { return type == 0 ? 8 : (type == 1 ? 4 : 2); }

// Build one group from a per-tile list of source columns.  When 'vdirect'
// is set the group's factors are the FP64 level storage itself (no copy):
// V and U column-major in place, one run of z columns per tile (rank padded
// to 16).  Otherwise both are copied and converted, in two layouts:
//   * U (row-major) and z by (tile dimension, row cluster): the columns of
//     all tiles that share rows form one run, padded to 16 once, so phase 2
//     takes one entry per run and leaf (or group of leaves) instead of one
//     per tile, however few columns each tile has in this group;
//   * V by (tile dimension, column cluster): the columns of all tiles that
//     read the same right-hand-side rows are contiguous, so a phase-1 item
//     of 64 columns spans tiles and one load of those rows serves all of
//     them; a per-column map gives each V column its z column.
// stage timings of the plan builds on stderr when HMX_LOG is set
struct StageLog {
    const char* label;
    bool on = std::getenv("HMX_LOG") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    explicit StageLog(const char* l) : label(l) {}
    void mark(Context& c, const char* what) {
        if (!on) return;
        HMX_HIP(hipStreamSynchronize(c.stream));
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[hmx] %s: %s %.1f ms\n", label, what,
                     std::chrono::duration<double, std::milli>(now - t).count());
        t = now;
    }
};

struct GroupTile {
    TileRef ref;
    std::vector<int> cols;  // source column indices in the tile
};

void build_group(Context& c, const HMatrix& h, int S, bool grouped, MfGroup& G, int type,
                 const std::vector<GroupTile>& tiles, bool vdirect, const std::vector<float>* su,
                 const std::vector<float>* sv, const std::vector<int64_t>* colBase, float zmul = 1.f,
                 int split = 0) {
    StageLog lg("plan group");
    HMX_CHECK(!split || (type == 1 && !vdirect && su), HMX_ESTATE, "split V: FP32 groups with column scales");
    const bool udirect = vdirect;
    const int b = (int)h.b;
    const size_t es = tsize(type);
    const size_t esu = split == 2 ? sizeof(_Float16) : es;  // split U: offsets into the hi copy
    G.type = type;
    const int T = (int)tiles.size();
    auto cnt_of = [&](int t) { return (int)tiles[t].cols.size(); };
    // FP32 / FP16 copies: every tile's columns padded to a multiple of 4 (zero
    // columns of U and V, zero scales) in both orders, for phase 1's vector
    // stores of z
    const bool pad4 = !vdirect && type != 0;
    const bool vfrag = !vdirect;  // V copies in fragment order (urm_index)
    auto cnt_pad = [&](int t) { return pad4 ? (cnt_of(t) + 3) / 4 * 4 : cnt_of(t); };
    auto scales = [&](int t, int q, float& fu, float& fv) {
        fu = fv = 1.f;
        if (su) {
            const int j = tiles[t].cols[q];
            fu = (*su)[(*colBase)[t] + j];
            fv = (*sv)[(*colBase)[t] + j];
        }
    };
    // first padded row of a tile's row / column cluster in the plan's space
    auto row_start = [&](const TileRef& R) { return R.leafBase * b + (int64_t)R.outCl * R.m; };
    auto col_start = [&](const TileRef& R) { return R.leafBase * b + (int64_t)R.inCl * R.m; };
    auto sorted_by = [&](bool byRow) {
        std::vector<int> ord;
        for (int t = 0; t < T; ++t)
            if (cnt_of(t)) ord.push_back(t);
        if (!vdirect)
            std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
                const TileRef &A = tiles[x].ref, &B = tiles[y].ref;
                const int64_t ca = byRow ? row_start(A) : col_start(A), cb = byRow ? row_start(B) : col_start(B);
                return A.m != B.m ? A.m > B.m : ca < cb;
            });
        return ord;
    };
    std::vector<MfItem> items;
    std::vector<int64_t> redCol, colPart;
    std::vector<int32_t> colNpart, zmap;
    std::vector<float> colScale;
    std::vector<std::vector<MfEntry>> leafE(S);
    // grouped (64-row leaves): phase 2 runs one workgroup per 4 leaves, and
    // runs spanning all 4 are listed once per group
    std::vector<std::vector<MfEntry>> grpE(grouped ? S / 4 : 0);
    std::vector<CvCol> vcols, ucols;
    std::vector<TrTile> trt;
    {
        size_t cols = 0;
        for (int t = 0; t < T; ++t) cols += (size_t)cnt_pad(t);
        if (!vdirect) {
            ucols.reserve(cols);
            vcols.reserve(cols);
        }
        colScale.reserve(cols + cols / 4);
        zmap.reserve(cols);
    }
    std::vector<int64_t> zb(T, -1);  // z column of each tile's first column in this group
    int64_t zcol = 0, slot = 0, vElems = 0, uElems = 0;
    uint64_t fbytes = 0;
    auto add_entries = [&](int m, int64_t leaf0, const MfEntry& proto, const double* ucm) {
        // U at the first row of the leaf (or group) inside the run's U;
        // leaf0 = the run's first leaf in the plan's leaf space
        const int bus = m / b;
        auto at = [&](int r0) {
            MfEntry e = proto;
            e.u = ucm ? (const void*)(ucm + (size_t)r0)
                      : (const void*)(uintptr_t)(((size_t)(uintptr_t)proto.u + (size_t)r0 * proto.rpad) * esu);
            return e;
        };
        if (grouped && bus >= 4)
            for (int q = 0; q < bus / 4; ++q) grpE[leaf0 / 4 + q].push_back(at(q * 4 * b));
        else
            for (int q = 0; q < bus; ++q) leafE[leaf0 + q].push_back(at(q * b));
    };
    // ---- U and z ----
    {
        const std::vector<int> ord = sorted_by(true);
        size_t i = 0;
        while (i < ord.size()) {
            const TileRef& R0 = tiles[ord[i]].ref;
            size_t j = i + 1;
            if (!udirect)
                while (j < ord.size() && tiles[ord[j]].ref.m == R0.m && row_start(tiles[ord[j]].ref) == row_start(R0))
                    ++j;
            const int m = R0.m;
            int cntRun = 0;
            for (size_t k = i; k < j; ++k) cntRun += cnt_pad(ord[k]);
            const int rpad = (cntRun + kPad - 1) / kPad * kPad;
            const int64_t cd0 = (int64_t)ucols.size();
            for (size_t k = i; k < j; ++k) {
                const int t = ord[k];
                const TileRef& R = tiles[t].ref;
                const double* Usrc = R.lv->U.p + (size_t)R.lv->h_offsets[R.t] * m;
                zb[t] = zcol + (int64_t)(ucols.size() - cd0);
                for (int q = 0; q < cnt_of(t); ++q) {
                    float fu, fv;
                    scales(t, q, fu, fv);
                    ucols.push_back(CvCol{Usrc + (size_t)tiles[t].cols[q] * m, 0, m, fu, 0, 0});
                    colScale.push_back(fu * fv * zmul);
                }
                for (int q = cnt_of(t); q < cnt_pad(t); ++q) {
                    ucols.push_back(CvCol{nullptr, 0, m, 1.f, 0, 0});
                    colScale.push_back(0.f);
                }
            }
            for (int q = cntRun; q < rpad; ++q) colScale.push_back(0.f);
            const int64_t leaf0 = row_start(R0) / b;
            if (udirect) {  // one tile, U column-major in place (all its columns, in order)
                const TileRef& R = R0;
                add_entries(m, leaf0, MfEntry{nullptr, rpad, (int32_t)zcol, cntRun, m},
                            R.lv->U.p + (size_t)R.lv->h_offsets[R.t] * m);
            } else {
                for (int r0 = 0; r0 < m; r0 += 32)
                    for (int c0 = 0; c0 < rpad; c0 += 32) trt.push_back(TrTile{cd0, uElems, m, cntRun, rpad, r0, c0, 0});
                add_entries(m, leaf0, MfEntry{(const void*)(uintptr_t)uElems, rpad, (int32_t)zcol, cntRun, 0},
                            nullptr);
                uElems += (int64_t)m * rpad;
            }
            zcol += rpad;
            HMX_CHECK(zcol < ((int64_t)1 << 31), HMX_EUNSUPPORTED, "matrix-core plan: more than 2^31 columns in one group");
            i = j;
        }
    }
    lg.mark(c, "U runs (host)");
    // ---- V and the phase-1 items ----
    // items over group columns [k0, k1), all of tile dimension m and column
    // cluster inCl; V at vptr (in place) or element offset vOff (copy)
    auto emit_items = [&](int64_t k0, int64_t k1, int m, int64_t colStart, const char* vptr, int64_t vOff) {
        const int itemRows = type == 0 ? kMfRows : mixed_item_rows((int64_t)S * b);
        const size_t esv = split ? sizeof(_Float16) : es;  // split: offsets into the hi copy
        const int nch = (m + itemRows - 1) / itemRows;
        const int rows = std::min(m, itemRows);
        const bool partial = nch > 1 || !h.planDirectZ;
        const int64_t cnt = k1 - k0;
        for (int64_t c0 = 0; c0 < cnt; c0 += kMfCols)
            for (int q = 0; q < nch; ++q) {
                MfItem it;
                // fragment-ordered copies: the row chunk q starts 16*q*itemRows
                // into its column block (urm_index(c0, q*itemRows, m))
                const size_t e = (size_t)c0 * m + (size_t)q * itemRows * (vfrag ? 16 : 1);
                it.v = vptr ? (const void*)(vptr + e * es) : (const void*)(uintptr_t)(((size_t)vOff + e) * esv);
                it.xoff = (int32_t)(colStart + q * itemRows);
                it.rows = rows;
                it.ld = m;
                it.ncols = (int32_t)std::min<int64_t>(kMfCols, cnt - c0);
                it.out = partial ? slot + c0 * nch + q : -1;
                it.zidx = k0 + c0;
                it.nch = nch;
                it.direct = partial ? 0 : 1;
                items.push_back(it);
            }
        if (partial) {
            for (int64_t k = k0; k < k1; ++k) {
                redCol.push_back(zmap[k]);
                colPart.push_back(slot + (k - k0) * nch);
                colNpart.push_back(nch);
            }
            slot += cnt * nch;
        }
    };
    {
        const std::vector<int> ord = sorted_by(false);
        size_t i = 0;
        while (i < ord.size()) {
            const TileRef& R0 = tiles[ord[i]].ref;
            size_t j = i + 1;
            if (!vdirect)
                while (j < ord.size() && tiles[ord[j]].ref.m == R0.m && col_start(tiles[ord[j]].ref) == col_start(R0))
                    ++j;
            const int m = R0.m;
            const int64_t k0 = (int64_t)zmap.size(), v0 = vElems;
            int32_t runCol = 0;
            // FP64 copies column by column; FP32/FP16 copies in fragment order
            // from the run's first element (dst = v0, column runCol)
            auto vcol = [&](const double* src, float scale) {
                vcols.push_back(vfrag ? CvCol{src, v0, m, scale, runCol, 0} : CvCol{src, vElems, m, scale, 0, 0});
                ++runCol;
                vElems += m;
            };
            for (size_t k = i; k < j; ++k) {
                const int t = ord[k];
                const TileRef& R = tiles[t].ref;
                const double* Vsrc = R.lv->V.p + (size_t)R.lv->h_offsets[R.t] * m;
                for (int q = 0; q < cnt_of(t); ++q) {
                    float fu, fv;
                    scales(t, q, fu, fv);
                    if (!vdirect) vcol(Vsrc + (size_t)tiles[t].cols[q] * m, fv);
                    else vElems += m;
                    zmap.push_back((int32_t)(zb[t] + q));
                }
                for (int q = cnt_of(t); q < cnt_pad(t); ++q) {
                    vcol(nullptr, 1.f);
                    zmap.push_back((int32_t)(zb[t] + q));
                }
                fbytes += 2ull * (uint64_t)cnt_pad(t) * m * es;
            }
            // a fragment-ordered run takes whole 16-column blocks (the
            // columns past its last are never read)
            if (vfrag) vElems = v0 + (int64_t)(runCol + 15) / 16 * 16 * m;
            if (vdirect)
                emit_items(k0, (int64_t)zmap.size(), m, col_start(R0),
                           (const char*)(R0.lv->V.p + (size_t)R0.lv->h_offsets[R0.t] * m), 0);
            else
                emit_items(k0, (int64_t)zmap.size(), m, col_start(R0), nullptr, v0);
            i = j;
        }
    }
    lg.mark(c, "V runs and items (host)");
    // items that read the same right-hand-side block (same tile dimension,
    // column cluster and row chunk) next to each other (already so for the
    // copied layout)
    if (h.planLocality)
        std::stable_sort(items.begin(), items.end(), [](const MfItem& a, const MfItem& b2) {
            return a.ld != b2.ld ? a.ld > b2.ld : a.xoff < b2.xoff;
        });
    G.vfrag = vfrag;
    G.split = split;
    G.vLoOff = split ? vElems : 0;  // the lo copies follow the hi copies (2 x 2 bytes per element)
    G.uLoOff = split == 2 ? uElems : 0;
    G.numColsPad = zcol;
    G.numPart = slot;
    G.factorBytes = fbytes;
    // storage
    if (!vdirect) G.V.alloc(std::max<int64_t>(1, vElems) * es);
    if (!udirect) G.Urm.alloc(std::max<int64_t>(1, uElems) * es);
    for (auto& it : items)
        if (!vdirect) it.v = (const void*)(G.V.p + (uintptr_t)it.v);
    std::vector<int32_t> leafPtr(S + 1, 0);
    std::vector<MfEntry> ents;
    for (int leaf = 0; leaf < S; ++leaf) {
        leafPtr[leaf] = (int32_t)ents.size();
        for (MfEntry e : leafE[leaf]) {
            if (!udirect) e.u = (const void*)(G.Urm.p + (uintptr_t)e.u);
            ents.push_back(e);
        }
    }
    leafPtr[S] = (int32_t)ents.size();
    G.grouped = grouped;
    if (grouped) {
        std::vector<int32_t> gptr(S / 4 + 1, 0);
        std::vector<MfEntry> gents;
        for (int g = 0; g < S / 4; ++g) {
            gptr[g] = (int32_t)gents.size();
            for (MfEntry e : grpE[g]) {
                if (!udirect) e.u = (const void*)(G.Urm.p + (uintptr_t)e.u);
                gents.push_back(e);
            }
        }
        gptr[S / 4] = (int32_t)gents.size();
        upload(c, G.grpPtr, gptr);
        upload(c, G.grpEntries, gents);
    }
    G.numItems = (int32_t)items.size();
    upload(c, G.items, items);
    upload(c, G.zmap, zmap);
    G.numRed = (int64_t)redCol.size();
    upload(c, G.redCol, redCol);
    upload(c, G.colPart, colPart);
    upload(c, G.colNpart, colNpart);
    upload(c, G.leafPtr, leafPtr);
    upload(c, G.entries, ents);
    if (su) {
        upload(c, G.colScale, colScale);
        std::vector<float> colScaleV(zmap.size());
        for (size_t k = 0; k < zmap.size(); ++k) colScaleV[k] = colScale[zmap[k]];
        upload(c, G.colScaleV, colScaleV);
    }
    lg.mark(c, "entries and uploads");
    DBuf<CvCol> d_u, d_v;
    DBuf<TrTile> d_t;
    upload(c, d_u, ucols);
    upload(c, d_v, vcols);
    upload(c, d_t, trt);
    lg.mark(c, "column descriptor uploads");
    if (!vcols.empty()) {
        const unsigned g = (unsigned)((vcols.size() + 3) / 4);
        if (type == 0) k_cvt_cols<double><<<g, 256, 0, c.stream>>>((int64_t)vcols.size(), d_v.p, (double*)G.V.p);
        else if (split) k_cvt_cols_split<<<g, 256, 0, c.stream>>>((int64_t)vcols.size(), d_v.p, (_Float16*)G.V.p, vElems);
        else if (type == 1) k_cvt_cols<float><<<g, 256, 0, c.stream>>>((int64_t)vcols.size(), d_v.p, (float*)G.V.p);
        else k_cvt_cols<_Float16><<<g, 256, 0, c.stream>>>((int64_t)vcols.size(), d_v.p, (_Float16*)G.V.p);
        HMX_LAUNCHED();
    }
    if (!trt.empty()) {
        const unsigned g = (unsigned)trt.size();
        if (type == 0) k_cvt_rowmajor<double><<<g, 256, 0, c.stream>>>((int64_t)trt.size(), d_t.p, d_u.p, (double*)G.Urm.p);
        else if (split == 2)
            k_cvt_rowmajor_split<<<g, 256, 0, c.stream>>>((int64_t)trt.size(), d_t.p, d_u.p, (_Float16*)G.Urm.p, uElems);
        else if (type == 1) k_cvt_rowmajor<float><<<g, 256, 0, c.stream>>>((int64_t)trt.size(), d_t.p, d_u.p, (float*)G.Urm.p);
        else k_cvt_rowmajor<_Float16><<<g, 256, 0, c.stream>>>((int64_t)trt.size(), d_t.p, d_u.p, (_Float16*)G.Urm.p);
        HMX_LAUNCHED();
    }
    HMX_HIP(hipStreamSynchronize(c.stream));
    lg.mark(c, "conversion kernels");
}

void check_bucket(const HMatrix& h) {
    HMX_CHECK(h.b == 32 || h.b == 64, HMX_EUNSUPPORTED, "the multi-RHS matvec kernels are built for bucketSize 32 or 64");
}

// leaf-level tiles of rank above b/2 are applied as dense blocks
bool is_near(const HMatrix& h, const TileRef& R) {
    return h.planDenseNear && R.m == (int)h.b && 2 * R.r > (int)h.b;
}

template <class T>
void build_near(Context& c, const HMatrix& h, int S, MfPlan& P, const std::vector<TileRef>& tiles) {
    const int b = (int)h.b;
    // blocks stored by row leaf (a leaf's blocks are one contiguous stretch,
    // which its phase-2 wave streams), in tile order within a leaf
    std::vector<std::vector<const TileRef*>> perLeaf(S);
    for (const TileRef& R : tiles)
        if (is_near(h, R)) perLeaf[R.leafBase + R.outCl].push_back(&R);
    std::vector<NearSrc> src;
    std::vector<int32_t> ptr(S + 1, 0);
    std::vector<int32_t> colLeaf;
    for (int leaf = 0; leaf < S; ++leaf) {
        ptr[leaf] = (int32_t)src.size();
        for (const TileRef* R : perLeaf[leaf]) {
            const double* U = R->lv->U.p + (size_t)R->lv->h_offsets[R->t] * R->m;
            const double* V = R->lv->V.p + (size_t)R->lv->h_offsets[R->t] * R->m;
            colLeaf.push_back((int32_t)(R->leafBase + R->inCl));
            src.push_back(NearSrc{U, V, (int64_t)src.size() * b * b, R->r, 0});
        }
    }
    ptr[S] = (int32_t)src.size();
    P.numNear = (int64_t)src.size();
    P.nearStore.alloc((size_t)std::max<int64_t>(1, P.numNear) * b * b * sizeof(T));
    if (P.numNear) {
        DBuf<NearSrc> d_src;
        upload(c, d_src, src);
        k_near_dense<T><<<(unsigned)P.numNear, 256, 0, c.stream>>>(b, d_src.p, (T*)P.nearStore.p);
        HMX_LAUNCHED();
        HMX_HIP(hipStreamSynchronize(c.stream));
    }
    std::vector<MfNear> list;
    for (int64_t q = 0; q < P.numNear; ++q)
        list.push_back(MfNear{(const void*)((const T*)P.nearStore.p + q * b * b), colLeaf[q], 0});
    upload(c, P.nearPtr, ptr);
    upload(c, P.near, list);
    HMX_HIP(hipStreamSynchronize(c.stream));
}

// The near-field block list of build_near (leaf-level tiles of rank above
// b/2, by row leaf) without the blocks themselves: for plans that evaluate
// them from the points.
void build_near_list(Context& c, const HMatrix& h, int S, MfPlan& P, const std::vector<TileRef>& tiles) {
    std::vector<std::vector<int32_t>> perLeaf(S);
    for (const TileRef& R : tiles)
        if (is_near(h, R)) perLeaf[R.leafBase + R.outCl].push_back((int32_t)(R.leafBase + R.inCl));
    std::vector<int32_t> ptr(S + 1, 0);
    std::vector<MfNear> list;
    for (int leaf = 0; leaf < S; ++leaf) {
        ptr[leaf] = (int32_t)list.size();
        for (int32_t cl : perLeaf[leaf]) list.push_back(MfNear{nullptr, cl, 0});
    }
    ptr[S] = (int32_t)list.size();
    P.numNear = (int64_t)list.size();
    upload(c, P.nearPtr, ptr);
    upload(c, P.near, list);
    // the same blocks merged per group of 4 leaves (one workgroup of the
    // phase-2 kernel), diagonals included: one entry per column leaf in
    // increasing order with the mask of the group's leaves that need it, so
    // the four waves walk one sequence and read each neighbour's right-hand
    // sides at about the same time
    const int G = (S + 3) / 4;
    std::vector<int32_t> gptr(G + 1, 0);
    std::vector<MfNear> glist;
    std::vector<std::pair<int32_t, int32_t>> tmp;
    for (int g = 0; g < G; ++g) {
        gptr[g] = (int32_t)glist.size();
        tmp.clear();
        for (int w = 0; w < 4 && 4 * g + w < S; ++w) {
            const int leaf = 4 * g + w;
            tmp.emplace_back(leaf, 1 << w);
            for (int32_t cl : perLeaf[leaf]) tmp.emplace_back(cl, 1 << w);
        }
        std::sort(tmp.begin(), tmp.end());
        for (size_t q = 0; q < tmp.size(); ++q) {
            if (!glist.empty() && q > 0 && tmp[q].first == tmp[q - 1].first) glist.back().pad |= tmp[q].second;
            else glist.push_back(MfNear{nullptr, tmp[q].first, tmp[q].second});
        }
    }
    gptr[G] = (int32_t)glist.size();
    upload(c, P.gnearPtr, gptr);
    upload(c, P.gnear, glist);
    HMX_HIP(hipStreamSynchronize(c.stream));
}

// FP64 matrix-core plan over one or several H-matrices (a batch is their
// block-diagonal operator): instance i owns leaves [leafBase_i, leafBase_i +
// S_i) of one padded leaf space.  Vectors come in as n_i consecutive entries
// per instance; with partial leaves they are gathered into the padded space
// and the product scattered back.
void build_fp64_plan_set(Context& c, const std::vector<HMatrix*>& hs, MfPlan& P) {
    HMX_CHECK(!hs.empty(), HMX_EINVAL, "empty plan set");
    const HMatrix& h0 = *hs[0];
    check_bucket(h0);
    const int b = (int)h0.b;
    std::vector<TileRef> all;
    std::vector<int32_t> gm;
    int64_t leafBase = 0, vecBase = 0;
    bool grouped = b == 64;
    for (HMatrix* h : hs) {
        HMX_CHECK(h->b == (uint32_t)b, HMX_EINVAL, "all instances of a batch need the same bucketSize");
        HMX_CHECK(h->leafOffsets.size() == (size_t)h->S + 1, HMX_ESTATE, "H-matrix without leaf offsets");
        std::vector<TileRef> t = all_tiles(*h, leafBase);
        all.insert(all.end(), t.begin(), t.end());
        std::vector<int32_t> g = padding_map(*h, vecBase);
        gm.insert(gm.end(), g.begin(), g.end());
        grouped = grouped && h->S % 4 == 0;
        leafBase += h->S;
        vecBase += h->n;
    }
    HMX_CHECK(leafBase * b < (int64_t)INT32_MAX, HMX_EUNSUPPORTED, "plan set too large for 32-bit row offsets");
    const int S = (int)leafBase;
    P.S = S;
    P.b = b;
    P.n = vecBase;
    P.npad = leafBase * b;
    P.gather = P.n != P.npad;
    if (P.gather) upload(c, P.gmap, gm);
    build_near<double>(c, h0, S, P, all);
    std::vector<GroupTile> tiles;
    for (const TileRef& R : all) {
        if (is_near(h0, R)) continue;
        GroupTile gt{R, {}};
        gt.cols.resize(R.r);
        std::iota(gt.cols.begin(), gt.cols.end(), 0);
        tiles.push_back(std::move(gt));
    }
    // factors in place (column-major, one phase-1 item run and one phase-2
    // entry per tile) or copied into the run layout (V by column cluster, U
    // row-major by row cluster: fewer, longer items and entries, whole-vector
    // U loads) when the copy fits
    bool copy = h0.planFp64Layout == 2;
    if (h0.planFp64Layout == 0) {
        uint64_t est = 0;
        for (const GroupTile& gt : tiles) est += 2ull * (uint64_t)(gt.ref.r + kPad) * gt.ref.m * sizeof(double);
        size_t fr = 0, totMem = 0;
        HMX_HIP(hipMemGetInfo(&fr, &totMem));
        copy = est <= fr / 2;
    }
    P.ucm = !copy;
    build_group(c, h0, S, grouped, P.g[0], 0, tiles, P.ucm, nullptr, nullptr, nullptr);
    const int64_t tot = (int64_t)S * b * b;
    P.Drm.alloc(tot * sizeof(double));
    int64_t off = 0;
    for (HMatrix* h : hs) {
        const int64_t e = (int64_t)h->S * b * b;
        k_diag_rowmajor<double><<<(unsigned)((e + 255) / 256), 256, 0, c.stream>>>(e, b, h->diag.p,
                                                                                  (double*)P.Drm.p + off);
        HMX_LAUNCHED();
        off += e;
    }
    HMX_HIP(hipStreamSynchronize(c.stream));
    P.dtype = 0;
    P.bytes = P.g[0].factorBytes + (uint64_t)tot * 8 + (uint64_t)P.numNear * b * b * 8;
    P.built = true;
}

// zLoOff >= 0 (T = FP16): z as FP16 hi + lo, lo at zLoOff
template <class ACC, class T>
void reduce_group(Context& c, MfGroup& G, int NR, const float* rhsScale = nullptr, int64_t zLoOff = -1) {
    if (G.numRed == 0) return;
    const unsigned blocks = (unsigned)((G.numRed + 3) / 4);
    const float* cs = G.colScale.n ? G.colScale.p : nullptr;
    auto launch = [&](auto k) {
        k<<<blocks, 256, 0, c.stream>>>(G.numRed, G.numColsPad, G.redCol.p, G.colPart.p, G.colNpart.p, cs,
                                        (const ACC*)G.zpart.p, (T*)G.zt.p, rhsScale, zLoOff);
    };
    if constexpr (sizeof(T) == 2) {
        if (zLoOff >= 0) {
            if (NR == 64) launch(k_mf_reduce<T, ACC, 64, true>);
            else launch(k_mf_reduce<T, ACC, 16, true>);
            HMX_LAUNCHED();
            return;
        }
    }
    if (NR == 64) launch(k_mf_reduce<T, ACC, 64>);
    else launch(k_mf_reduce<T, ACC, 16>);
    HMX_LAUNCHED();
}

// Workspaces.  The padding columns of z (rank rounded up to 16) are never
// written by the kernels and must hold zeros: U is zero there, but 0 * NaN
// left by an earlier call would not be.  Their positions depend on NR
// (zt_index), so z is cleared again whenever the pass width changes.
// Returns true when a workspace was allocated or cleared (launch sequences
// captured before then point at memory in another state).
bool ensure_ws(Context& c, MfGroup& G, int NR, size_t accBytes) {
    bool changed = false;
    if (G.wsNR < (size_t)NR) {
        G.zpart.alloc((size_t)std::max<int64_t>(1, G.numPart) * NR * accBytes);
        G.zt.alloc((size_t)std::max<int64_t>(1, G.numColsPad) * NR * tsize(G.type));
        G.wsNR = NR;
        G.ztNR = 0;
        changed = true;
    }
    if (G.ztNR != NR) {
        HMX_HIP(hipMemsetAsync(G.zt.p, 0, G.zt.bytes(), c.stream));
        G.ztNR = NR;
        changed = true;
    }
    return changed;
}

// workgroups per XCD chunk in phase 1 (0: hardware order)
int phase1_chunk() {
    static const int v = [] {
        const char* e = std::getenv("HMX_P1_CHUNK");
        return e ? std::max(0, std::atoi(e)) : 16;
    }();
    return v;
}

template <class T, int CG, int NB>
void phase1(Context& c, MfGroup& G, const T* Xb, int sh, int kvalid, int xcdChunk = -1) {
    if (G.numItems == 0) return;
    const float* cs = G.colScaleV.n ? G.colScaleV.p : nullptr;
    const int chunk = xcdChunk >= 0 ? xcdChunk : phase1_chunk();
    unsigned blocks = (unsigned)((G.numItems + 3) / 4);
    if (chunk > 0) blocks = (blocks + 8 * chunk - 1) / (8 * chunk) * (8 * chunk);
    // FP64 groups read V in place (column-major) or from a fragment-ordered
    // copy: separate instantiations, so the in-place kernel's code is not
    // touched by the other path (one wave per SIMD, sensitive to it)
    auto k1 = G.vfrag ? k_mf_phase1<T, CG, NB, true> : k_mf_phase1<T, CG, NB, false>;
    k1<<<blocks, 256, 0, c.stream>>>(G.items.p, G.numItems, chunk, Xb, sh, (typename Mfma<T>::acc_t*)G.zpart.p,
                                     (T*)G.zt.p, G.zmap.p, cs, kvalid, nullptr, 0, nullptr, 0);
    HMX_LAUNCHED();
}

// the lo half of a split group's FP16 z (MfGroup::split == 2): after the hi
// half, whose size is fixed by the widest pass the workspace was made for
int64_t split_zlo(const MfGroup& G) { return std::max<int64_t>(1, G.numColsPad) * (int64_t)G.wsNR; }

// phase 1 of a mixed plan's FP32 group held as FP16 hi + lo (MfGroup::split):
// TZ = float writes FP32 z in true units (rhsScale), TZ = _Float16 FP16 hi +
// lo z in right-hand-side units
template <int NB, class TZ>
void phase1_split(Context& c, MfGroup& G, const _Float16* Xh, const _Float16* Xl, int sh, int kvalid,
                  const float* rhsScale) {
    if (G.numItems == 0) return;
    HMX_CHECK(G.split == (sizeof(TZ) == 2 ? 2 : 1) && G.vfrag && G.colScaleV.n && Xl, HMX_ESTATE,
              "split phase 1: plan not built for it");
    const int chunk = phase1_chunk();
    unsigned blocks = (unsigned)((G.numItems + 3) / 4);
    if (chunk > 0) blocks = (blocks + 8 * chunk - 1) / (8 * chunk) * (8 * chunk);
    k_mf_phase1<_Float16, 4, NB, true, true, TZ><<<blocks, 256, 0, c.stream>>>(
        G.items.p, G.numItems, chunk, Xh, sh, (float*)G.zpart.p, (TZ*)G.zt.p, G.zmap.p, G.colScaleV.p, kvalid, Xl,
        G.vLoOff, rhsScale, sizeof(TZ) == 2 ? split_zlo(G) : 0);
    HMX_LAUNCHED();
}

struct Events {
    hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
    bool on = false;
    void create() {
        for (auto& x : e) HMX_HIP(hipEventCreate(&x));
        on = true;
    }
    void rec(Context& c, int i) {
        if (on) HMX_HIP(hipEventRecord(e[i], c.stream));
    }
    ~Events() {
        for (auto& x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

// One pass of up to 16*NB right-hand sides (kk of them real) through an FP64
// plan: X and Y hold the caller's entries (ld ldx / ldy); with partial leaves
// X is gathered into the padded space and the product scattered back.
template <int NB>
void run_fp64(Context& c, MfPlan& P, const double* X, int64_t ldx, int kk, double* Y, int64_t ldy, Events& ev) {
    MfGroup& G = P.g[0];
    const int NR = 16 * NB, sh = P.b == 64 ? 6 : 5;
    const double* Xb = (const double*)P.Xb.p;
    ev.rec(c, 0);
    const int64_t tot = P.npad * NR;
    k_x_blocked<double, double><<<(unsigned)((tot + 255) / 256), 256, 0, c.stream>>>(
        tot, P.npad, NR, sh, X, ldx, kk, P.gather ? P.gmap.p : nullptr, nullptr, (double*)P.Xb.p);
    HMX_LAUNCHED();
    phase1<double, 4, NB>(c, G, Xb, sh, kk, P.xcdChunk);  // 4 column groups = the 64 columns of an item
    ev.rec(c, 1);
    reduce_group<double, double>(c, G, NR);
    ev.rec(c, 2);
    double* Yo = P.gather ? (double*)P.Yp.p : Y;
    const int ldo = (int)(P.gather ? P.npad : ldy);
    {
        const int* gp = G.grouped ? G.grpPtr.p : nullptr;
        const MfEntry* ge = G.grouped ? G.grpEntries.p : nullptr;
        auto k2 = P.b == 64 ? (P.ucm ? k_mf_phase2g<64, true, double, double, double, NB, double>
                                     : k_mf_phase2g<64, false, double, double, double, NB, double>)
                            : (P.ucm ? k_mf_phase2g<32, true, double, double, double, NB, double>
                                     : k_mf_phase2g<32, false, double, double, double, NB, double>);
        k2<<<(P.S + 3) / 4, 256, 0, c.stream>>>(P.S, G.leafPtr.p, G.entries.p, gp, ge, (const double*)G.zt.p,
                                                nullptr, nullptr, nullptr, nullptr, nullptr, (const double*)P.Drm.p,
                                                P.nearPtr.p, P.near.p, Xb, Yo, ldo, kk, nullptr, 1.f, nullptr, nullptr,
                                                0.f, nullptr, nullptr, 0, 0, 1.f);
    }
    HMX_LAUNCHED();
    if (P.gather) {
        k_y_scatter<double><<<(unsigned)((P.npad * kk + 255) / 256), 256, 0, c.stream>>>(P.npad, kk, P.gmap.p,
                                                                                        (const double*)P.Yp.p, Y, ldy);
        HMX_LAUNCHED();
    }
    ev.rec(c, 3);
}

// One pass of up to 16*NB right-hand sides (kk of them real) through the
// mixed plan, X and Y FP32 with the caller's entries (ld ldx / ldy); with
// partial leaves X is gathered into the padded space and Y scattered back.
template <int NB>
void run_mixed(Context& c, MfPlan& P, const float* X, int64_t ldx, int kk, float* Y, int64_t ldy, Events& ev) {
    MfGroup& Ghi = P.g[0];
    MfGroup& Glo = P.g[1];
    // right-hand sides in blocks of 16 bytes of rows: 8 (FP16), 4 (FP32)
    const int NR = 16 * NB, sh16 = 3, sh32 = 2;
    const int32_t* gm = P.gather ? P.gmap.p : nullptr;
    ev.rec(c, 0);
    // right-hand sides to FP16 with one scale per column
    k_colmaxabs_part<<<dim3(kXSlices, kk), 256, 0, c.stream>>>((int)P.n, X, ldx, P.xpart.p);
    HMX_LAUNCHED();
    k_colmaxabs_fin<<<kk, 64, 0, c.stream>>>(kXSlices, P.xpart.p, P.xscale.p);
    HMX_LAUNCHED();
    HMX_CHECK(P.npad % 32 == 0 && NR <= 64, HMX_EUNSUPPORTED, "mixed plan: padded space not in 32-row chunks");
    k_x_mixed<<<(unsigned)(P.npad / 32), 256, 0, c.stream>>>(P.npad, NR, X, ldx, kk, gm, P.xscale.p, (float*)P.Xb.p,
                                                            P.X16.p, P.X16lo.n ? P.X16lo.p : nullptr);
    HMX_LAUNCHED();
    const float* Xb = (const float*)P.Xb.p;
    if (Ghi.split == 2) phase1_split<NB, _Float16>(c, Ghi, P.X16.p, P.X16lo.p, sh16, kk, nullptr);
    else if (Ghi.split == 1) phase1_split<NB, float>(c, Ghi, P.X16.p, P.X16lo.p, sh16, kk, P.xscale.p);
    else phase1<float, 4, NB>(c, Ghi, Xb, sh32, kk);
    phase1<_Float16, 4, NB>(c, Glo, P.X16.p, sh16, kk);
    ev.rec(c, 1);
    if (Ghi.split == 2) reduce_group<float, _Float16>(c, Ghi, NR, nullptr, split_zlo(Ghi));
    else reduce_group<float, float>(c, Ghi, NR, Ghi.split ? P.xscale.p : nullptr);
    reduce_group<float, _Float16>(c, Glo, NR);
    ev.rec(c, 2);
    float* Yo = P.gather ? (float*)P.Yp.p : Y;
    const int ldo = (int)(P.gather ? P.npad : ldy);
    {
        const bool gr = Ghi.grouped && Glo.grouped;
        const bool spa = Ghi.split == 2;
        // a split group's z is FP16 hi + lo (the kernel's ZA), else FP32
        auto launch = [&](auto k2, const auto* za) {
            k2<<<(P.S + 3) / 4, 256, 0, c.stream>>>(
                P.S, Ghi.leafPtr.p, Ghi.entries.p, gr ? Ghi.grpPtr.p : nullptr, gr ? Ghi.grpEntries.p : nullptr, za,
                Glo.leafPtr.p, Glo.entries.p, gr ? Glo.grpPtr.p : nullptr, gr ? Glo.grpEntries.p : nullptr,
                (const _Float16*)Glo.zt.p, (const float*)P.Drm.p, P.nearOtf ? P.gnearPtr.p : P.nearPtr.p,
                P.nearOtf ? P.gnear.p : P.near.p, Xb, Yo, ldo, kk, P.xscale.p, P.loMul,
                P.nearOtf ? reinterpret_cast<const float4*>(P.ptsL.p) : nullptr,
                reinterpret_cast<const double4*>(P.origin.p), P.nlog2e, P.X16.p, P.X16lo.p, spa ? Ghi.uLoOff : 0,
                spa ? split_zlo(Ghi) : 0, P.hiMul);
        };
        const _Float16* zh = (const _Float16*)Ghi.zt.p;
        const float* zf = (const float*)Ghi.zt.p;
        if (spa && P.b == 64) launch(k_mf_phase2g<64, false, float, _Float16, _Float16, NB, float, true>, zh);
        else if (spa) launch(k_mf_phase2g<32, false, float, _Float16, _Float16, NB, float, true>, zh);
        else if (P.b == 64) launch(k_mf_phase2g<64, false, float, float, _Float16, NB, float>, zf);
        else launch(k_mf_phase2g<32, false, float, float, _Float16, NB, float>, zf);
    }
    HMX_LAUNCHED();
    if (P.gather) {
        k_y_scatter<float><<<(unsigned)((P.npad * kk + 255) / 256), 256, 0, c.stream>>>(P.npad, kk, P.gmap.p,
                                                                                       (const float*)P.Yp.p, Y, ldy);
        HMX_LAUNCHED();
    }
    ev.rec(c, 3);
}

void finish_timing(Events& ev, float* acc) {
    if (!ev.on) return;
    HMX_HIP(hipEventSynchronize(ev.e[3]));
    float t;
    for (int q = 0; q < 3; ++q) {
        HMX_HIP(hipEventElapsedTime(&t, ev.e[q], ev.e[q + 1]));
        acc[q] += t;
    }
    HMX_HIP(hipEventElapsedTime(&t, ev.e[0], ev.e[3]));
    acc[3] += t;
}

}  // namespace

// Workspaces of an FP64 plan for passes of NR right-hand sides
bool ensure_fp64_ws(Context& c, MfPlan& P, int NR) {
    bool changed = ensure_ws(c, P.g[0], NR, sizeof(double));
    const size_t xb = (size_t)P.npad * NR * sizeof(double);
    if (P.Xb.n < xb) { P.Xb.alloc(xb); changed = true; }
    if (P.gather && P.Yp.n < xb) { P.Yp.alloc(xb); changed = true; }
    return changed;
}

// Y = H X for k right-hand sides (ld = the plan's n) in passes of 64 (k a
// multiple of 64) or 16, the last pass padded with zero columns
void run_fp64_plan(Context& c, MfPlan& P, uint32_t k, const double* d_in, double* d_out, Events& ev) {
    const int NB = (k % 64 == 0) ? 4 : 1, NR = 16 * NB;
    for (uint32_t s0 = 0; s0 < k; s0 += NR) {
        const int kk = (int)std::min<uint32_t>(NR, k - s0);
        const double* X = d_in + (size_t)s0 * P.n;
        double* Y = d_out + (size_t)s0 * P.n;
        if (NB == 4) run_fp64<4>(c, P, X, P.n, kk, Y, P.n, ev);
        else run_fp64<1>(c, P, X, P.n, kk, Y, P.n, ev);
    }
}

void hmatvec_multi(Context& c, HMatrix& h, uint32_t k, const double* d_in, double* d_out, float* ms, int iters) {
    HMX_CHECK(k >= 1, HMX_EINVAL, "multi-RHS matvec: vectorWidth must be >= 1");
    HMX_CHECK(d_in && d_out && d_in != d_out, HMX_EINVAL, "bad vector pointers");
    if (!h.mf64.built) build_fp64_plan_set(c, {&h}, h.mf64);
    MfPlan& P = h.mf64;
    ensure_fp64_ws(c, P, (k % 64 == 0) ? 64 : 16);
    Events ev;
    if (ms) ev.create();
    const int reps = ms ? std::max(1, iters) : 1;
    float acc[4] = {0, 0, 0, 0};
    for (int it = 0; it < reps; ++it) {
        run_fp64_plan(c, P, k, d_in, d_out, ev);
        finish_timing(ev, acc);
    }
    if (ms)
        for (int q = 0; q < 4; ++q) ms[q] = acc[q] / reps;
}

bool batch_mf_stale(const Batch& B) {
    if (!B.mf.built) return false;
    for (size_t i = 0; i < B.hs.size(); ++i)
        if (B.hs[i]->factorGen != B.mfGen[i]) return true;
    return false;
}

// the FP64 plan of a batch, rebuilt when an instance's factors were
// overwritten since (the plan may hold copies of them); true when the plan
// was rebuilt or a workspace allocated or cleared
bool batch_mf_prepare(Context& c, Batch& B, uint32_t k) {
    bool changed = false;
    if (batch_mf_stale(B)) {
        const int xc = B.mf.xcdChunk;
        B.mf = MfPlan();
        B.mf.xcdChunk = xc;
    }
    if (!B.mf.built) {
        build_fp64_plan_set(c, B.hs, B.mf);
        B.mfGen.clear();
        for (const HMatrix* h : B.hs) B.mfGen.push_back(h->factorGen);
        changed = true;
    }
    return ensure_fp64_ws(c, B.mf, (k % 64 == 0) ? 64 : 16) || changed;
}

void batch_mf_run(Context& c, Batch& B, uint32_t k, const double* d_in, double* d_out) {
    Events ev;
    run_fp64_plan(c, B.mf, k, d_in, d_out, ev);
}

int default_xcd_chunk() { return phase1_chunk(); }

void hmatrix_make_mixed(Context& c, HMatrix& h, double target) {
    StageLog lg("mixed plan");
    check_bucket(h);
    HMX_CHECK(h.leafOffsets.size() == (size_t)h.S + 1, HMX_ESTATE, "H-matrix without leaf offsets");
    HMX_CHECK(target > 0 && target < 1, HMX_EINVAL, "target relative error must be in (0, 1)");
    MfPlan& P = h.mfMixed;
    P = MfPlan();
    std::vector<TileRef> tiles = all_tiles(h);
    // per-column statistics
    std::vector<ColDesc> cd;
    std::vector<int64_t> colBase;
    {
        size_t cols = 0;
        for (const TileRef& R : tiles) cols += (size_t)R.r;
        cd.reserve(cols);
        colBase.reserve(tiles.size());
    }
    for (const TileRef& R : tiles) {
        colBase.push_back((int64_t)cd.size());
        const double* U = R.lv->U.p + (size_t)R.lv->h_offsets[R.t] * R.m;
        const double* V = R.lv->V.p + (size_t)R.lv->h_offsets[R.t] * R.m;
        for (int j = 0; j < R.r; ++j) cd.push_back(ColDesc{U + (size_t)j * R.m, V + (size_t)j * R.m, R.m, 0});
    }
    const int64_t nc = (int64_t)cd.size();
    std::vector<double> v2(nc), um(nc), vm(nc);
    if (nc) {
        DBuf<ColDesc> d_cd;
        upload(c, d_cd, cd);
        DBuf<double> d_v2(nc), d_um(nc), d_vm(nc);
        k_colstats<<<(unsigned)((nc + 3) / 4), 256, 0, c.stream>>>(nc, d_cd.p, d_v2.p, d_um.p, d_vm.p);
        HMX_LAUNCHED();
        HMX_HIP(hipMemcpyAsync(v2.data(), d_v2.p, nc * 8, hipMemcpyDeviceToHost, c.stream));
        HMX_HIP(hipMemcpyAsync(um.data(), d_um.p, nc * 8, hipMemcpyDeviceToHost, c.stream));
        HMX_HIP(hipMemcpyAsync(vm.data(), d_vm.p, nc * 8, hipMemcpyDeviceToHost, c.stream));
        HMX_HIP(hipStreamSynchronize(c.stream));
    }
    lg.mark(c, "column statistics");
    // ||H||_F^2 = sum of column energies (one factor of every tile has
    // orthonormal columns) + dense blocks
    std::vector<double> hd(h.diag.n);
    HMX_HIP(hipMemcpy(hd.data(), h.diag.p, h.diag.bytes(), hipMemcpyDeviceToHost));
    double frob = 0;
    for (double x : hd) frob += x * x;
    for (double x : v2) frob += x;
    // partial leaves: the diagonal blocks hold 1 at each padding position
    frob -= (double)((int64_t)h.S * h.b - (int64_t)h.n);
    // FP16 storage error of a column: rounding of u, v, x and z (rms 2^-11/sqrt(3) each)
    const double delta = std::ldexp(1.0, -11) / std::sqrt(3.0);
    // aimed at 0.9 of the target: the budget sets the expected error, and a
    // small plan's error scatters around it (3.004e-6 at a 3e-6 target, n =
    // 5001, 5 right-hand sides, without the margin)
    const double aim = 0.9 * target;
    const double budget = (aim / (2.0 * delta)) * (aim / (2.0 * delta)) * frob;
    // columns of tiles applied as dense blocks stay out of the selection
    std::vector<char> dense(nc, 0);
    for (size_t t = 0; t < tiles.size(); ++t)
        if (is_near(h, tiles[t]))
            for (int j = 0; j < tiles[t].r; ++j) dense[colBase[t] + j] = 1;
    // The columns in increasing energy order while the sum stays within the
    // budget, without sorting all of them: the bit patterns of non-negative
    // doubles order like their values, so a histogram over the top 16 bits
    // (count and energy per bucket) gives the bucket where the running sum
    // crosses the budget; the buckets below it are taken whole and only that
    // bucket is sorted (by energy, then index) and taken greedily.
    std::vector<char> lo(nc, 0);
    double used = 0;
    {
        auto key = [&](int64_t j) {
            uint64_t u;
            std::memcpy(&u, &v2[j], 8);
            return (uint32_t)(u >> 48);
        };
        std::vector<double> bsum(1 << 16, 0.0);
        for (int64_t j = 0; j < nc; ++j)
            if (!dense[j]) bsum[key(j)] += v2[j];
        uint32_t cut = 1u << 16;  // first bucket not taken whole
        for (uint32_t k = 0; k < (1u << 16); ++k) {
            if (used + bsum[k] > budget) {
                cut = k;
                break;
            }
            used += bsum[k];
        }
        std::vector<std::pair<double, int64_t>> edge;
        for (int64_t j = 0; j < nc; ++j) {
            if (dense[j]) continue;
            const uint32_t k = key(j);
            if (k < cut) lo[j] = 1;
            else if (k == cut) edge.emplace_back(v2[j], j);
        }
        std::sort(edge.begin(), edge.end());
        for (const auto& e : edge) {
            if (used + e.first > budget) break;
            used += e.first;
            lo[e.second] = 1;
        }
    }
    lg.mark(c, "FP16 column selection");
    std::vector<float> su(nc), sv(nc);
    for (int64_t j = 0; j < nc; ++j) {
        su[j] = um[j] > 0 ? (float)um[j] : 1.f;
        sv[j] = vm[j] > 0 ? (float)vm[j] : 1.f;
    }
    // FP16 z of column j = (V16_j . X16) * su_j * sv_j * 2^E with |V16|, |X16|
    // <= 1, so |z| <= m_j * su_j * sv_j * 2^E: one exponent E for the plan
    // puts the largest such bound at 2^15 (below the FP16 maximum 65504, and
    // the small columns as far above the subnormal range as that allows).
    // Phase 2 multiplies the FP16 sum by 2^-E and the right-hand-side scale.
    double zb = 0;
    for (size_t t = 0; t < tiles.size(); ++t)
        for (int j = 0; j < tiles[t].r; ++j)
            if (lo[colBase[t] + j])
                zb = std::max(zb, (double)su[colBase[t] + j] * (double)sv[colBase[t] + j] * tiles[t].m);
    const int zexp = zb > 0 ? (int)std::floor(std::log2(32768.0 / zb)) : 0;
    const bool grouped = h.b == 64 && h.S >= 4;
    // with the points at hand (direct assembly) the diagonal and near-field
    // blocks are evaluated in phase 2 instead of stored
    P.nearOtf = h.pts.n >= (size_t)h.S * h.b * 4 && h.ell > 0;
    if (P.nearOtf) {
        build_near_list(c, h, (int)h.S, P, tiles);
        const int64_t npad = (int64_t)h.S * h.b;
        P.ptsL.alloc((size_t)npad * 4);
        P.origin.alloc((size_t)h.S * 4);
        k_local_points<<<(unsigned)((npad + 255) / 256), 256, 0, c.stream>>>(
            npad, (int)h.b, reinterpret_cast<const double4*>(h.pts.p), reinterpret_cast<float4*>(P.ptsL.p),
            reinterpret_cast<double4*>(P.origin.p));
        HMX_LAUNCHED();
        P.nlog2e = (float)(-1.4426950408889634 / h.ell);
    } else {
        build_near<float>(c, h, (int)h.S, P, tiles);
    }
    lg.mark(c, "near field");
    std::vector<GroupTile> gh, gl;
    int64_t hi = 0, lc = 0;
    for (size_t t = 0; t < tiles.size(); ++t) {
        GroupTile a{tiles[t], {}}, b2{tiles[t], {}};
        if (is_near(h, tiles[t])) {  // keep the group lists aligned with colBase (empty entries)
            gh.push_back(std::move(a));
            gl.push_back(std::move(b2));
            continue;
        }
        for (int j = 0; j < tiles[t].r; ++j) (lo[colBase[t] + j] ? b2.cols : a.cols).push_back(j);
        hi += (int64_t)a.cols.size();
        lc += (int64_t)b2.cols.size();
        gh.push_back(std::move(a));
        gl.push_back(std::move(b2));
    }
    lg.mark(c, "column split");
    {
        // The FP32 group as FP16 hi + lo scaled per column (HMX_MIXED_SPLIT:
        // 1, the default, V only, U and z FP32; 2 V, U and z, both phases on
        // the FP16 matrix cores; 0 FP32 copies).  Mode 2's z holds V^T x in
        // right-hand-side units times su_j sv_j 2^Ea, bounded by
        // m_j su_j sv_j 2^Ea <= 2^15 as for the FP16 group; phase 2 brings
        // the group's sum to the FP16 group's units with 2^(E - Ea).  The one
        // exponent leaves the group's small columns with z's lo part in the
        // FP16 subnormals (about 17 bits), an error outside the energy
        // budget: N = 2^20 3D at a 1e-5 target, 64 right-hand sides, 7.1e-6
        // instead of 5.9e-6 (64-RHS call 10.7 instead of 11.2 ms), so it is
        // not the default.
        const char* e = std::getenv("HMX_MIXED_SPLIT");
        const int split = e ? std::min(2, std::max(0, std::atoi(e))) : 1;
        if (split == 2) {
            double zbA = 0;
            for (size_t t = 0; t < gh.size(); ++t)
                for (int j : gh[t].cols)
                    zbA = std::max(zbA, (double)su[colBase[t] + j] * (double)sv[colBase[t] + j] * tiles[t].m);
            const int ea = zbA > 0 ? (int)std::floor(std::log2(32768.0 / zbA)) : 0;
            build_group(c, h, (int)h.S, grouped, P.g[0], 1, gh, false, &su, &sv, &colBase, (float)std::ldexp(1.0, ea),
                        2);
            P.hiMul = (float)std::ldexp(1.0, zexp - ea);
        } else if (split == 1) {
            const std::vector<float> ones(su.size(), 1.f);
            build_group(c, h, (int)h.S, grouped, P.g[0], 1, gh, false, &ones, &sv, &colBase, 1.f, 1);
        } else {
            build_group(c, h, (int)h.S, grouped, P.g[0], 1, gh, false, nullptr, nullptr, nullptr);
        }
    }
    lg.mark(c, "FP32 group");
    build_group(c, h, (int)h.S, grouped, P.g[1], 2, gl, false, &su, &sv, &colBase, (float)std::ldexp(1.0, zexp));
    lg.mark(c, "FP16 group");
    const int64_t tot = (int64_t)h.S * h.b * h.b;
    if (!P.nearOtf) {
        P.Drm.alloc(tot * sizeof(float));
        k_diag_rowmajor<float><<<(unsigned)((tot + 255) / 256), 256, 0, c.stream>>>(tot, (int)h.b, h.diag.p,
                                                                                   (float*)P.Drm.p);
        HMX_LAUNCHED();
    }
    HMX_HIP(hipStreamSynchronize(c.stream));
    P.dtype = 1;
    P.loMul = (float)std::ldexp(1.0, -zexp);
    P.S = (int)h.S;
    P.b = (int)h.b;
    P.n = h.n;
    P.npad = (int64_t)h.S * h.b;
    P.gather = P.n != P.npad;
    if (P.gather) upload(c, P.gmap, padding_map(h, 0));
    HMX_HIP(hipStreamSynchronize(c.stream));
    P.frobH2 = frob;
    P.loEnergy = used;
    P.target = target;
    P.hiCols = hi;
    P.loCols = lc;
    // bytes read per call: factors, FP16 (and split FP32) column scales, and the stored dense
    // blocks (none when they are evaluated on the fly: then the points)
    P.bytes = P.g[0].factorBytes + P.g[1].factorBytes + (uint64_t)lc * 4 + (P.g[0].split ? (uint64_t)hi * 4 : 0) +
              (P.nearOtf ? (uint64_t)h.S * h.b * 16 + (uint64_t)h.S * 32
                         : (uint64_t)tot * 4 + (uint64_t)P.numNear * h.b * h.b * 4);
    P.built = true;
}

// Workspaces of the mixed plan for passes of NR right-hand sides; true when
// any was allocated or cleared
bool ensure_mixed_ws(Context& c, MfPlan& P, int NR) {
    bool changed = ensure_ws(c, P.g[0], NR, sizeof(float));
    changed = ensure_ws(c, P.g[1], NR, sizeof(float)) || changed;
    if (P.x16n < (size_t)P.npad * NR) {
        P.X16.alloc((size_t)P.npad * NR);
        if (P.nearOtf || P.g[0].split) P.X16lo.alloc((size_t)P.npad * NR);
        P.Xb.alloc((size_t)P.npad * NR * sizeof(float));
        if (P.gather) P.Yp.alloc((size_t)P.npad * NR * sizeof(float));
        P.x16n = (size_t)P.npad * NR;
        changed = true;
    }
    if (P.xscale.n < (size_t)NR) {
        P.xscale.alloc(NR);
        P.xpart.alloc((size_t)NR * kXSlices);
        // scales past a call's width are never set; the split FP32 group's
        // reduction multiplies its (zero) partials of those columns by them
        HMX_HIP(hipMemsetAsync(P.xscale.p, 0, P.xscale.bytes(), c.stream));
        changed = true;
    }
    return changed;
}

// Y = H X for any k >= 1 (FP32, ld n): passes of 64 (k a multiple of 64) or
// 16, the last one padded with zero columns
void hmatvec_mixed(Context& c, HMatrix& h, uint32_t k, const float* d_in, float* d_out, float* ms, int iters) {
    HMX_CHECK(h.mfMixed.built, HMX_ESTATE, "mixed-precision factors not built (call hmx_hmatrix_make_mixed)");
    HMX_CHECK(k >= 1, HMX_EINVAL, "mixed matvec: vectorWidth must be >= 1");
    HMX_CHECK(d_in && d_out && d_in != d_out, HMX_EINVAL, "bad vector pointers");
    MfPlan& P = h.mfMixed;
    const int NB = (k % 64 == 0) ? 4 : 1;
    const int NR = 16 * NB;
    ensure_mixed_ws(c, P, NR);
    Events ev;
    if (ms) ev.create();
    const int reps = ms ? std::max(1, iters) : 1;
    float acc[4] = {0, 0, 0, 0};
    for (int it = 0; it < reps; ++it) {
        for (uint32_t s0 = 0; s0 < k; s0 += NR) {
            const int kk = (int)std::min<uint32_t>(NR, k - s0);
            const float* X = d_in + (size_t)s0 * h.n;
            float* Y = d_out + (size_t)s0 * h.n;
            if (NB == 4) run_mixed<4>(c, P, X, h.n, kk, Y, h.n, ev);
            else run_mixed<1>(c, P, X, h.n, kk, Y, h.n, ev);
        }
        finish_timing(ev, acc);
    }
    if (ms)
        for (int q = 0; q < 4; ++q) ms[q] = acc[q] / reps;
}

}  // namespace hmx
