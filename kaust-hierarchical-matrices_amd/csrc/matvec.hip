// H-matrix x vector (HMatrixVecMult, src/HMatrixVectorMultiplication.cu:10-26;
// cutlassDiagonalXVec, src/cutlassDiagonalXVector.cu:54-76; cutlassHierarchicalXVec,
// src/cutlassHMatrixXVector.cu:330-398).
//
// y = D x + sum over all levels and tiles t of U_t (V_t^T x_col(t)), for any
// admissibility structure, in three launches (all levels at once):
//   phase 1  z_t = V_t^T x : one wave per work item, a contiguous run of V
//            read with 16-byte loads; long columns are split into chunks
//            (2048 elements by default, tunable) that produce partial sums,
//            short columns are packed many per item and reduced with segmented
//            wave reductions;
//   reduce   z = sum of the chunk partials (fixed order -> deterministic);
//   phase 2  one workgroup per leaf (32 or 64 rows): y_leaf = D_leaf x_leaf + sum of the
//            leaf's rows of every U_t covering it times z_t; with 64-row leaves
//            the tiles of 128 rows and more are left to a second pass, one
//            workgroup per 128-row block adding them to y in a fixed order.
//            No atomics.
// The factors stay where the build left them (no per-call V transpose as in
// src/cutlassHMatrixXVector.cu:118-150, no per-call host problem assembly).
#include <algorithm>

#include "hmx_device.h"
#include "hmx_internal.h"

namespace hmx {

constexpr int kRowBlock = 128;  // rows per workgroup of the tall-tile phase-2 pass

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// factor data is read once per call: non-temporal loads keep x and z in L2
typedef double f64x2_t __attribute__((ext_vector_type(2)));
__device__ inline double2 ld_nt2(const double* __restrict__ p) {
    const f64x2_t v = __builtin_nontemporal_load(reinterpret_cast<const f64x2_t*>(p));
    return make_double2(v.x, v.y);
}

template <int N>
__device__ inline void vload_d(const double* __restrict__ p, double (&out)[N]) {
    static_assert(N % 2 == 0, "pairs of doubles");
#pragma unroll
    for (int i = 0; i < N; i += 2) {
        const double2 v = ld_nt2(p + i);
        out[i] = v.x;
        out[i + 1] = v.y;
    }
}

// ---- phase 1 ---------------------------------------------------------------
// A wave walks its item in steps of 128 elements (2 per lane, 16-byte loads),
// U steps at a time: the V and x loads of all U steps are issued before any
// of their arithmetic, so every lane has U loads in flight.
template <int K>
__global__ void __launch_bounds__(256) k_mv_phase1(const MvItem* __restrict__ items, int numItems,
                                                   const double* __restrict__ X, int ldx,
                                                   double* __restrict__ zpart) {
    constexpr int U = K <= 4 ? 4 : (K == 8 ? 2 : 1);  // keep the register file for occupancy at wide K
    const int lane = threadIdx.x & 63;
    // wave-uniform item index (readfirstlane): scalar item loads, no exec
    // masks on the loops; 13.3-13.7 -> 12.5 ms at N = 2^20, one right-hand side
    const int wid = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    if (wid >= numItems) return;
    const MvItem it = items[wid];
    const int len = it.len;
    const int m = it.colLen ? it.colLen : len;  // column length inside this item (power of two)
    const double* __restrict__ xs = X + it.xoff;
    if (m == 128) {
        // one column per step, and a lane meets the same two x entries in every step
        double2 xr[K];
#pragma unroll
        for (int s = 0; s < K; ++s) xr[s] = *reinterpret_cast<const double2*>(xs + (size_t)s * ldx + 2 * lane);
        int col = 0;
        for (int e0 = 0; e0 < len; e0 += 128 * U) {
            double2 vv[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                vv[u] = e0 + 128 * u < len ? ld_nt2(it.v + e0 + 128 * u + 2 * lane)
                                           : make_double2(0.0, 0.0);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (e0 + 128 * u >= len) break;
#pragma unroll
                for (int s = 0; s < K; ++s) {
                    const double r = wave_sum(fma(vv[u].y, xr[s].y, vv[u].x * xr[s].x));
                    if (lane == 0) zpart[(size_t)(it.out + col) * K + s] = r;
                }
                ++col;
            }
        }
    } else if (m > 128) {
        // every 128-element step lies inside one column
        double acc[K];
#pragma unroll
        for (int s = 0; s < K; ++s) acc[s] = 0.0;
        int col = 0;
        for (int e0 = 0; e0 < len; e0 += 128 * U) {
            double2 vv[U], xx[U][K];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = e0 + 128 * u + 2 * lane;
                const bool ok = e0 + 128 * u < len;  // uniform: len is a multiple of 128
                vv[u] = ok ? ld_nt2(it.v + e) : make_double2(0.0, 0.0);
                const int xi = e & (m - 1);
#pragma unroll
                for (int s = 0; s < K; ++s)
                    xx[u][s] = ok ? *reinterpret_cast<const double2*>(xs + (size_t)s * ldx + xi) : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (e0 + 128 * u >= len) break;
#pragma unroll
                for (int s = 0; s < K; ++s) {
                    acc[s] = fma(vv[u].x, xx[u][s].x, acc[s]);
                    acc[s] = fma(vv[u].y, xx[u][s].y, acc[s]);
                }
                if (((e0 + 128 * u + 128) & (m - 1)) == 0) {
#pragma unroll
                    for (int s = 0; s < K; ++s) {
                        const double r = wave_sum(acc[s]);
                        if (lane == 0) zpart[(size_t)(it.out + col) * K + s] = r;
                        acc[s] = 0.0;
                    }
                    ++col;
                }
            }
        }
    } else {
        // m in {32, 64}: one step covers 128/m columns; m/2 lanes per column.
        // A lane meets the same two x entries in every step: load them once.
        const int g = m >> 1;
        const int xi = (2 * lane) & (m - 1);
        double2 xr[K];
#pragma unroll
        for (int s = 0; s < K; ++s) xr[s] = *reinterpret_cast<const double2*>(xs + (size_t)s * ldx + xi);
        for (int e0 = 0; e0 < len; e0 += 128 * U) {
            double2 vv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = e0 + 128 * u + 2 * lane;
                vv[u] = e < len ? ld_nt2(it.v + e) : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (e0 + 128 * u >= len) break;
                const int e = e0 + 128 * u + 2 * lane;
                double p[K];
#pragma unroll
                for (int s = 0; s < K; ++s) p[s] = fma(vv[u].y, xr[s].y, vv[u].x * xr[s].x);
                for (int off = 1; off < g; off <<= 1)
#pragma unroll
                    for (int s = 0; s < K; ++s) p[s] += __shfl_xor(p[s], off, 64);
                if (e < len && (lane & (g - 1)) == 0) {
                    const int col = e / m;
#pragma unroll
                    for (int s = 0; s < K; ++s) zpart[(size_t)(it.out + col) * K + s] = p[s];
                }
            }
        }
    }
}

// ---- reduce partials ---------------------------------------------------------
__global__ void __launch_bounds__(256) k_mv_reduce(int64_t numCols, int K, const int64_t* __restrict__ colPart,
                                                   const int32_t* __restrict__ colNpart,
                                                   const double* __restrict__ zpart, double* __restrict__ z) {
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= numCols * K) return;
    int64_t g = e / K;
    int s = (int)(e % K);
    const int64_t p0 = colPart[g];
    const int np = colNpart[g];
    double v = 0.0;
    for (int c = 0; c < np; ++c) v += zpart[(p0 + c) * K + s];
    z[e] = v;
}

// ---- phase 2 -----------------------------------------------------------------
// 256 threads = 16 column groups x 16 lanes; a lane owns RPL = BL/16
// consecutive rows of the leaf (BL = bucketSize) and a column group walks
// columns cg, cg+16, ...
template <int BL, int K>
__global__ void __launch_bounds__(256) k_mv_phase2(const int32_t* __restrict__ leafPtr,
                                                   const MvEntry* __restrict__ entries,
                                                   const double* __restrict__ diag, const double* __restrict__ z,
                                                   const double* __restrict__ X, int ldx, double* __restrict__ Y,
                                                   const int32_t* __restrict__ nearPtr,
                                                   const int32_t* __restrict__ nearCol,
                                                   const double* __restrict__ nearD) {
    constexpr int RPL = BL / 16;
    static_assert(RPL == 2 || RPL == 4, "bucketSize 32 or 64");
    __shared__ double red[4][BL * K];
    const int leaf = blockIdx.x;
    const int l16 = threadIdx.x & 15, cg = threadIdx.x >> 4;
    // lane l16 holds rows 2*l16 + {0, 1} (+ 32 for its second pair with
    // 64-row leaves): each 16-byte load instruction of a column reads 256
    // contiguous bytes (a lane reading 4 consecutive rows as two loads used
    // half of every line per instruction)
    const int r0 = 2 * l16;
    auto rowq = [&](int q) { return (q >> 1) * 32 + r0 + (q & 1); };
    auto load_col = [&](const double* __restrict__ colp, double (&out)[RPL]) {
        double h[2];
        vload_d<2>(colp + r0, h);
        out[0] = h[0];
        out[1] = h[1];
        if constexpr (RPL == 4) {
            vload_d<2>(colp + 32 + r0, h);
            out[2] = h[0];
            out[3] = h[1];
        }
    };
    double acc[RPL][K];
#pragma unroll
    for (int q = 0; q < RPL; ++q)
#pragma unroll
        for (int s = 0; s < K; ++s) acc[q][s] = 0.0;
    // dense blocks (column-major): the diagonal block, then the leaf's
    // near-field blocks; columns c = cg, cg+16, ... of each
    const int q0 = nearPtr ? nearPtr[leaf] : 0, q1 = nearPtr ? nearPtr[leaf + 1] : 0;
    for (int q = q0 - 1; q < q1; ++q) {
        const double* D = q < q0 ? diag + (size_t)leaf * BL * BL : nearD + (size_t)q * BL * BL;
        const double* xl = X + (size_t)(q < q0 ? leaf : nearCol[q]) * BL;
#pragma unroll
        for (int cc = 0; cc < BL / 16; ++cc) {
            const int c = cg + 16 * cc;
            double dv[RPL];
            load_col(D + c * BL, dv);
#pragma unroll
            for (int s = 0; s < K; ++s) {
                const double xv = xl[(size_t)s * ldx + c];
#pragma unroll
                for (int q2 = 0; q2 < RPL; ++q2) acc[q2][s] = fma(dv[q2], xv, acc[q2][s]);
            }
        }
    }
    const int e0 = leafPtr[leaf], e1 = leafPtr[leaf + 1];
    for (int e = e0; e < e1; ++e) {
        const MvEntry en = entries[e];
        const double* __restrict__ U = en.u;
        const double* __restrict__ zz = z + (size_t)en.zbase * K;
        const int m = en.m;
        // 64 columns per step: a column group's four columns (cg, cg+16,
        // cg+32, cg+48) load together, the ones past the rank masked off, so
        // a tile of rank <= 64 costs one round trip instead of one per column
        for (int j0 = cg; j0 < en.rank; j0 += 64) {
            double u[4][RPL], zv[4][K];
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const int j = j0 + 16 * p;
                if (j < en.rank) {
                    load_col(U + (size_t)j * m, u[p]);
#pragma unroll
                    for (int s = 0; s < K; ++s) zv[p][s] = zz[(size_t)j * K + s];
                }
            }
#pragma unroll
            for (int p = 0; p < 4; ++p)
                if (j0 + 16 * p < en.rank)
#pragma unroll
                    for (int s = 0; s < K; ++s)
#pragma unroll
                        for (int q = 0; q < RPL; ++q) acc[q][s] = fma(u[p][q], zv[p][s], acc[q][s]);
        }
    }
    // fold the 4 column groups of a wave (lanes l16 + 16g), then the 4 waves
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < RPL; ++q)
#pragma unroll
        for (int s = 0; s < K; ++s) {
            acc[q][s] += __shfl_xor(acc[q][s], 16, 64);
            acc[q][s] += __shfl_xor(acc[q][s], 32, 64);
        }
    if (lane < 16)
#pragma unroll
        for (int s = 0; s < K; ++s)
#pragma unroll
            for (int q = 0; q < RPL; ++q) red[wave][s * BL + rowq(q)] = acc[q][s];
    __syncthreads();
    for (int o = threadIdx.x; o < BL * K; o += 256) {
        const double v = (red[0][o] + red[1][o]) + (red[2][o] + red[3][o]);
        const int s = o / BL, row = o % BL;
        Y[(size_t)s * ldx + (size_t)leaf * BL + row] = v;
    }
}

// ---- phase 2, tall tiles ------------------------------------------------------
// One workgroup per block of RB rows adds the tiles of dimension >= RB to y
// (after the per-leaf pass wrote it).  Every wave covers all RB rows, RPL =
// RB/64 consecutive rows per lane, so one column's segment is one contiguous
// RB*8-byte read (the per-leaf pass reads 512-byte pieces of the same columns
// at m*8-byte strides); wave w takes columns w, w+4, ..., CPS of them per
// step.  Blocks are numbered XCD-locally: the blocks of one tall tile share
// its z and meet in one L2.  The four waves' partial sums are folded in LDS
// in a fixed order.
template <int RB, int K, int CPS = 4>
__global__ void __launch_bounds__(256) k_mv_phase2b(const int32_t* __restrict__ blkPtr,
                                                    const MvEntry* __restrict__ entries,
                                                    const double* __restrict__ z, double* __restrict__ Y, int ldy) {
    constexpr int RPL = RB / 64;
    static_assert(RPL % 2 == 0, "16-byte loads");
    __shared__ double red[4][RB];
    const int blk = xcd_local(blockIdx.x, gridDim.x);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r0 = RPL * lane;
    double acc[RPL][K];
#pragma unroll
    for (int q = 0; q < RPL; ++q)
#pragma unroll
        for (int s = 0; s < K; ++s) acc[q][s] = 0.0;
    const int e0 = blkPtr[blk], e1 = blkPtr[blk + 1];
    for (int e = e0; e < e1; ++e) {
        const MvEntry en = entries[e];
        const double* __restrict__ U = en.u;
        const double* __restrict__ zz = z + (size_t)en.zbase * K;
        for (int j0 = w; j0 < en.rank; j0 += 4 * CPS) {
            double u[CPS][RPL], zv[CPS][K];
#pragma unroll
            for (int p = 0; p < CPS; ++p) {
                const int j = j0 + 4 * p;
                if (j < en.rank) {
                    vload_d<RPL>(U + (size_t)j * en.m + r0, u[p]);
#pragma unroll
                    for (int s = 0; s < K; ++s) zv[p][s] = zz[(size_t)j * K + s];
                }
            }
#pragma unroll
            for (int p = 0; p < CPS; ++p)
                if (j0 + 4 * p < en.rank)
#pragma unroll
                    for (int s = 0; s < K; ++s)
#pragma unroll
                        for (int q = 0; q < RPL; ++q) acc[q][s] = fma(u[p][q], zv[p][s], acc[q][s]);
        }
    }
    double* __restrict__ Yb = Y + (size_t)blk * RB;
#pragma unroll
    for (int s = 0; s < K; ++s) {
#pragma unroll
        for (int q = 0; q < RPL; ++q) red[w][r0 + q] = acc[q][s];
        __syncthreads();
        for (int row = threadIdx.x; row < RB; row += 256)
            Yb[(size_t)s * ldy + row] += (red[0][row] + red[1][row]) + (red[2][row] + red[3][row]);
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_transpose_blocks(int S, int b, const double* __restrict__ D,
                                                          double* __restrict__ DT) {
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int64_t per = (int64_t)b * b;
    if (e >= (int64_t)S * per) return;
    int64_t leaf = e / per;
    int r = (int)(e % per);
    int c = r / b, q = r % b;
    DT[leaf * per + (int64_t)q * b + c] = D[e];
}

// dense near-field blocks, column-major: out[q][col*b + row] = sum_a U[a*b + row] V[a*b + col]
__global__ void __launch_bounds__(256) k_near_dense_cm(int b, const NearRef* __restrict__ src,
                                                       double* __restrict__ out) {
    const NearRef t = src[blockIdx.x];
    double* o = out + (size_t)blockIdx.x * b * b;
    for (int e = threadIdx.x; e < b * b; e += 256) {
        const int col = e / b, row = e % b;
        double v = 0.0;
        for (int a = 0; a < t.r; ++a) v = fma(t.U[(size_t)a * b + row], t.V[(size_t)a * b + col], v);
        o[e] = v;
    }
}

// vectors (ldx apart) <-> the padded space (npad apart), K right-hand sides
__global__ void __launch_bounds__(256) k_gather(int64_t npad, int K, const int32_t* __restrict__ gmap,
                                                const double* __restrict__ X, int64_t ldx, double* __restrict__ Xp) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= npad * K) return;
    const int64_t s = e / npad, p = e % npad;
    const int g = gmap[p];
    Xp[e] = g >= 0 ? X[s * ldx + g] : 0.0;
}

__global__ void __launch_bounds__(256) k_scatter(int64_t npad, int K, const int32_t* __restrict__ gmap,
                                                 const double* __restrict__ Yp, double* __restrict__ Y, int64_t ldy) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= npad * K) return;
    const int64_t s = e / npad, p = e % npad;
    const int g = gmap[p];
    if (g >= 0) Y[s * ldy + g] = Yp[e];
}

// Work lists of one H-matrix: phase-1 items over the input-side factors,
// partial-slot bookkeeping, phase-2 entries per output leaf.  Positions are
// in the padded space (leaf*b + i).
PlanHost plan_host(const HMatrix& h, bool transpose, int chunk, int rowBlock, bool denseNear) {
    const int L = h.numLevels, b = (int)h.b, S = (int)h.S;
    HMX_CHECK(b == 32 || b == 64, HMX_EUNSUPPORTED, "HMatrixVecMult kernels are built for bucketSize 32 or 64");
    HMX_CHECK(chunk >= 128 && (chunk & (chunk - 1)) == 0, HMX_EINVAL, "phase-1 chunk must be a power of two >= 128");
    PlanHost H;
    H.b = b;
    std::vector<std::vector<MvEntry>> leafEntries(S);
    const int nblk = rowBlock > 0 ? (int)((int64_t)S * b / rowBlock) : 0;
    std::vector<std::vector<MvEntry>> blkEntries(nblk);
    int64_t slot = 0, g = 0;
    for (int l = 1; l < L; ++l) {
        const HLevel& lv = h.levels[l - 1];
        const std::vector<int32_t>& tiles = h.structure->tileIndices[l - 1];
        const int m = lv.m;
        const int bus = m / b;
        for (int t = 0; t < lv.numTiles; ++t) {
            const int r = lv.h_ranks[t];
            if (r == 0) continue;
            const uint32_t u = compact16((uint32_t)tiles[t]), v = compact16((uint32_t)tiles[t] >> 1);
            const uint32_t inCl = transpose ? v : u, outCl = transpose ? u : v;
            const double* Pin = (transpose ? lv.U.p : lv.V.p) + (size_t)lv.h_offsets[t] * m;
            const double* Pout = (transpose ? lv.V.p : lv.U.p) + (size_t)lv.h_offsets[t] * m;
            if (denseNear && m == b && 2 * r > b) {  // fewer bytes as a dense block, and no z
                H.nears.push_back(NearRef{Pout, Pin, r, (int32_t)outCl, (int32_t)inCl, 0});
                continue;
            }
            H.lowRankBytes += 16ull * (uint64_t)r * (uint64_t)m;
            const int nch = m >= chunk ? m / chunk : 1;
            for (int j = 0; j < r; ++j) {
                H.colPart.push_back(slot + (int64_t)j * nch);
                H.colNpart.push_back(nch);
            }
            if (m >= chunk) {
                for (int j = 0; j < r; ++j)
                    for (int q = 0; q < nch; ++q)
                        H.items.push_back(MvItem{Pin + (size_t)j * m + (size_t)q * chunk, (int32_t)(inCl * m + q * chunk),
                                                 chunk, 0, (int32_t)(slot + (int64_t)j * nch + q)});
            } else {
                const int cpi = chunk / m;
                for (int j0 = 0; j0 < r; j0 += cpi) {
                    const int nc = std::min(cpi, r - j0);
                    H.items.push_back(MvItem{Pin + (size_t)j0 * m, (int32_t)(inCl * m), nc * m, m, (int32_t)(slot + j0)});
                }
            }
            if (rowBlock > 0 && m >= rowBlock) {
                const int nb = m / rowBlock;
                for (int q = 0; q < nb; ++q)
                    blkEntries[outCl * nb + q].push_back(MvEntry{Pout + (size_t)q * rowBlock, m, r, (int32_t)g, 0});
            } else {
                for (int q = 0; q < bus; ++q)
                    leafEntries[outCl * bus + q].push_back(MvEntry{Pout + (size_t)q * b, m, r, (int32_t)g, 0});
            }
            slot += (int64_t)r * nch;
            g += r;
        }
    }
    HMX_CHECK(slot < (int64_t)INT32_MAX && g < (int64_t)INT32_MAX, HMX_EUNSUPPORTED, "matvec plan too large");
    H.leafPtr.assign(S + 1, 0);
    for (int leaf = 0; leaf < S; ++leaf) {
        H.leafPtr[leaf] = (int32_t)H.entries.size();
        H.entries.insert(H.entries.end(), leafEntries[leaf].begin(), leafEntries[leaf].end());
    }
    H.leafPtr[S] = (int32_t)H.entries.size();
    H.rowBlock = nblk > 0 ? rowBlock : 0;
    if (nblk > 0) {
        H.blkPtr.assign(nblk + 1, 0);
        for (int k = 0; k < nblk; ++k) {
            H.blkPtr[k] = (int32_t)H.blkEntries.size();
            H.blkEntries.insert(H.blkEntries.end(), blkEntries[k].begin(), blkEntries[k].end());
        }
        H.blkPtr[nblk] = (int32_t)H.blkEntries.size();
    }
    H.numCols = g;
    H.numPart = slot;
    return H;
}

template <class T>
static void put(Context& c, DBuf<T>& d, const std::vector<T>& h) {
    d.alloc(std::max<size_t>(1, h.size()));
    if (!h.empty()) HMX_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, c.stream));
}

void upload_plan(Context& c, MatvecPlan& P, const PlanHost& H) {
    P.numItems = (int32_t)H.items.size();
    P.numCols = H.numCols;
    P.numPart = H.numPart;
    P.numLeaves = (int32_t)H.leafPtr.size() - 1;
    put(c, P.items, H.items);
    put(c, P.colPart, H.colPart);
    put(c, P.colNpart, H.colNpart);
    put(c, P.leafPtr, H.leafPtr);
    put(c, P.entries, H.entries);
    P.rowBlock = H.rowBlock;
    P.numBlocks = H.rowBlock ? (int32_t)H.blkPtr.size() - 1 : 0;
    if (H.rowBlock) {
        put(c, P.blkPtr, H.blkPtr);
        put(c, P.blkEntries, H.blkEntries);
    }
    // dense near-field blocks, grouped by output leaf
    const int S = P.numLeaves, b = H.b;
    std::vector<NearRef> nears = H.nears;
    std::stable_sort(nears.begin(), nears.end(), [](const NearRef& x, const NearRef& y) { return x.outLeaf < y.outLeaf; });
    std::vector<int32_t> ptr(S + 1, 0), col(nears.size());
    for (const NearRef& nr : nears) ptr[nr.outLeaf + 1]++;
    for (int l = 0; l < S; ++l) ptr[l + 1] += ptr[l];
    for (size_t q = 0; q < nears.size(); ++q) col[q] = nears[q].inLeaf;
    P.numNear = (int64_t)nears.size();
    put(c, P.nearPtr, ptr);
    put(c, P.nearCol, col);
    P.nearD.alloc((size_t)std::max<int64_t>(1, P.numNear) * b * b);
    if (P.numNear) {
        DBuf<NearRef> d_src;
        put(c, d_src, nears);
        k_near_dense_cm<<<(unsigned)P.numNear, 256, 0, c.stream>>>(b, d_src.p, P.nearD.p);
        HMX_LAUNCHED();
        HMX_HIP(hipStreamSynchronize(c.stream));
    }
    P.bytes = H.lowRankBytes + 8ull * (uint64_t)(S + P.numNear) * b * b;
    P.lowRankBytes = H.lowRankBytes;
    P.wsK = 0;  // workspaces are sized for the previous plan
    HMX_HIP(hipStreamSynchronize(c.stream));
}

// padded position -> index of the KD-ordered vector (-1 on padding)
std::vector<int32_t> padding_map(const HMatrix& h, int64_t base) {
    const int S = (int)h.S, b = (int)h.b;
    std::vector<int32_t> gm((size_t)S * b, -1);
    for (int leaf = 0; leaf < S; ++leaf) {
        const int a0 = h.leafOffsets[leaf], cnt = h.leafOffsets[leaf + 1] - a0;
        for (int i = 0; i < cnt; ++i) gm[(size_t)leaf * b + i] = (int32_t)(base + a0 + i);
    }
    return gm;
}

// 64-row leaves: tiles of 128 rows and more go through the row-block pass of
// phase 2 (measured at N = 2^20, 3D: phase 2 14.9 -> 13.0 ms; blocks of 256
// rows 14.4 ms, of 64 rows 14.9 ms).  32-row leaves and batches: per-leaf
// pass only.
int row_block(const HMatrix& h) {
    const int64_t npad = (int64_t)h.S * h.b;
    return h.b == 64 && npad >= kRowBlock && npad % kRowBlock == 0 ? kRowBlock : 0;
}

// denseNear: leaf-level tiles of rank above b/2 as dense blocks (the plans
// of the products); off for the factor-gradient plans, which need z of
// every tile
void build_matvec_plan(Context& c, HMatrix& h, bool transpose, MatvecPlan& P, bool denseNear) {
    const int b = (int)h.b, S = (int)h.S;
    HMX_CHECK(h.leafOffsets.size() == (size_t)S + 1, HMX_ESTATE, "H-matrix without leaf offsets");
    upload_plan(c, P, plan_host(h, transpose, P.chunk, row_block(h), denseNear));
    P.n = h.n;
    P.npad = (int64_t)S * b;
    P.gather = h.n != (uint32_t)P.npad;
    if (P.gather) {
        put(c, P.gmap, padding_map(h, 0));
        HMX_HIP(hipStreamSynchronize(c.stream));
    }
    if (transpose) {
        P.diagT.alloc(h.diag.n);
        int64_t tot = (int64_t)h.diag.n;
        k_transpose_blocks<<<(unsigned)((tot + 255) / 256), 256, 0, c.stream>>>(S, b, h.diag.p, P.diagT.p);
        HMX_LAUNCHED();
        P.diag = P.diagT.p;
    } else {
        P.diag = h.diag.p;
    }
    HMX_HIP(hipStreamSynchronize(c.stream));
    P.built = true;
}

template <int K>
static void run_phases(Context& c, const MatvecPlan& P, const MvWork& W, int b, const double* X, int64_t ldx, double* Y,
                       hipEvent_t* ev) {
    if (ev) HMX_HIP(hipEventRecord(ev[0], c.stream));
    if (P.numItems > 0) {
        k_mv_phase1<K><<<(unsigned)((P.numItems + 3) / 4), 256, 0, c.stream>>>(P.items.p, P.numItems, X, (int)ldx,
                                                                                 W.zpart);
        HMX_LAUNCHED();
    }
    if (ev) HMX_HIP(hipEventRecord(ev[1], c.stream));
    if (P.numCols > 0) {
        k_mv_reduce<<<(unsigned)((P.numCols * K + 255) / 256), 256, 0, c.stream>>>(P.numCols, K, P.colPart.p,
                                                                                   P.colNpart.p, W.zpart, W.z);
        HMX_LAUNCHED();
    }
    if (ev) HMX_HIP(hipEventRecord(ev[2], c.stream));
    const int32_t* near = P.numNear ? P.nearPtr.p : nullptr;
    if (b == 64)
        k_mv_phase2<64, K><<<(unsigned)P.numLeaves, 256, 0, c.stream>>>(P.leafPtr.p, P.entries.p, P.diag, W.z, X,
                                                                         (int)ldx, Y, near, P.nearCol.p, P.nearD.p);
    else
        k_mv_phase2<32, K><<<(unsigned)P.numLeaves, 256, 0, c.stream>>>(P.leafPtr.p, P.entries.p, P.diag, W.z, X,
                                                                         (int)ldx, Y, near, P.nearCol.p, P.nearD.p);
    HMX_LAUNCHED();
    if (P.numBlocks > 0) {  // after the per-leaf pass: it adds to Y
        HMX_CHECK(P.rowBlock == kRowBlock, HMX_ESTATE, "matvec plan with an unknown row block");
        k_mv_phase2b<kRowBlock, K><<<(unsigned)P.numBlocks, 256, 0, c.stream>>>(P.blkPtr.p, P.blkEntries.p, W.z,
                                                                                 Y, (int)ldx);
        HMX_LAUNCHED();
    }
    if (ev) HMX_HIP(hipEventRecord(ev[3], c.stream));
}

// Workspaces for passes of up to K right-hand sides.  Returns true when they
// were (re)allocated: launch sequences captured before point at the old ones.
bool ensure_matvec_ws(MatvecPlan& P, uint32_t K) {
    if (P.wsK >= K) return false;
    P.zpart.alloc((size_t)std::max<int64_t>(1, P.numPart) * K);
    P.z.alloc((size_t)std::max<int64_t>(1, P.numCols) * K);
    if (P.gather) {
        P.Xp.alloc((size_t)P.npad * K);
        P.Yp.alloc((size_t)P.npad * K);
    }
    P.wsK = K;
    return true;
}

// K = kernel width (1, 4, 8 or 16); k (a multiple of K) right-hand sides in
// passes of K.  With partial leaves every pass gathers into and scatters from
// the padded space.
template <int K>
static void run_k(Context& c, const MatvecPlan& P, const MvWork& W, int b, uint32_t k, const double* X, int64_t ldx,
                  double* Y, hipEvent_t* ev) {
    for (uint32_t s0 = 0; s0 < k; s0 += K) {
        const double* Xs = X + (size_t)s0 * ldx;
        double* Ys = Y + (size_t)s0 * ldx;
        if (!P.gather) {
            run_phases<K>(c, P, W, b, Xs, ldx, Ys, ev);
            continue;
        }
        const int64_t tot = P.npad * K;
        k_gather<<<(unsigned)((tot + 255) / 256), 256, 0, c.stream>>>(P.npad, K, P.gmap.p, Xs, ldx, W.Xp);
        HMX_LAUNCHED();
        run_phases<K>(c, P, W, b, W.Xp, P.npad, W.Yp, ev);
        k_scatter<<<(unsigned)((tot + 255) / 256), 256, 0, c.stream>>>(P.npad, K, P.gmap.p, W.Yp, Ys, ldx);
        HMX_LAUNCHED();
    }
}

uint32_t pass_width(uint32_t rem) { return rem >= 16 ? 16 : (rem >= 8 ? 8 : (rem >= 4 ? 4 : 1)); }

// Workspace elements (doubles) of passes of up to K right-hand sides:
// partial sums, z, and with partial leaves the padded copies of x and y.
size_t matvec_ws_elems(const MatvecPlan& P, uint32_t K) {
    return ((size_t)std::max<int64_t>(1, P.numPart) + (size_t)std::max<int64_t>(1, P.numCols) +
            (P.gather ? 2 * (size_t)P.npad : 0)) * K;
}

// Any k: the widest kernels first (k = 21 runs 16 + 4 + 1 right-hand sides),
// on the workspaces W (sized for passes of pass_width(k) right-hand sides).
void run_matvec_plan_ws(Context& c, const MatvecPlan& P, const MvWork& W, int b, uint32_t k, const double* X, int64_t ldx,
                        double* Y, hipEvent_t* ev) {
    for (uint32_t s0 = 0; s0 < k;) {
        const uint32_t K = pass_width(k - s0), cnt = (k - s0) / K * K;
        const double* Xs = X + (size_t)s0 * ldx;
        double* Ys = Y + (size_t)s0 * ldx;
        switch (K) {
            case 1: run_k<1>(c, P, W, b, cnt, Xs, ldx, Ys, ev); break;
            case 4: run_k<4>(c, P, W, b, cnt, Xs, ldx, Ys, ev); break;
            case 8: run_k<8>(c, P, W, b, cnt, Xs, ldx, Ys, ev); break;
            default: run_k<16>(c, P, W, b, cnt, Xs, ldx, Ys, ev); break;
        }
        s0 += cnt;
    }
}

// The same on the plan's own workspaces (grown here when too small).
void run_matvec_plan(Context& c, MatvecPlan& P, int b, uint32_t k, const double* X, int64_t ldx, double* Y,
                     hipEvent_t* ev) {
    ensure_matvec_ws(P, pass_width(k));
    const MvWork W{P.zpart.p, P.z.p, P.Xp.p, P.Yp.p};
    run_matvec_plan_ws(c, P, W, b, k, X, ldx, Y, ev);
}

void hmatvec(Context& c, HMatrix& h, uint32_t k, const double* d_in, double* d_out, bool transpose, float* ms,
             int iters) {
    HMX_CHECK(k >= 1, HMX_EINVAL, "vectorWidth must be >= 1");
    HMX_CHECK(d_in && d_out, HMX_EINVAL, "null vector pointer");
    HMX_CHECK(d_in != d_out, HMX_EINVAL, "in and out must not alias");
    MatvecPlan& P = transpose ? h.planT : h.plan;
    if (!P.built) build_matvec_plan(c, h, transpose, P, h.planDenseNear != 0);
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    const bool prof = ms != nullptr;
    if (prof)
        for (auto& e : ev) HMX_HIP(hipEventCreate(&e));
    const int reps = prof ? std::max(1, iters) : 1;
    float acc[4] = {0, 0, 0, 0};
    for (int it = 0; it < reps; ++it) {
        run_matvec_plan(c, P, (int)h.b, k, d_in, (int64_t)h.n, d_out, prof ? ev : nullptr);
        if (prof) {
            HMX_HIP(hipEventSynchronize(ev[3]));
            float t;
            for (int q = 0; q < 3; ++q) {
                HMX_HIP(hipEventElapsedTime(&t, ev[q], ev[q + 1]));
                acc[q] += t;
            }
            HMX_HIP(hipEventElapsedTime(&t, ev[0], ev[3]));
            acc[3] += t;
        }
    }
    if (prof) {
        for (int q = 0; q < 4; ++q) ms[q] = acc[q] / reps;
        for (auto& e : ev) (void)hipEventDestroy(e);
    }
}

// ---- many H-matrices in one set of launches ----------------------------------
// The plans of all instances are concatenated: instance i's padded vector
// space starts at padBase_i, its leaves at leafBase_i, its z columns and
// partial slots after those of the instances before it; its vectors start at
// vecBase_i = sum of the previous instances' n inside every right-hand side.
void build_batch_plan(Context& c, const std::vector<HMatrix*>& hs, int chunk, MatvecPlan& P, DBuf<double>& diag) {
    HMX_CHECK(!hs.empty(), HMX_EINVAL, "empty batch");
    const int b = (int)hs[0]->b;
    PlanHost B;
    B.b = b;
    B.leafPtr.push_back(0);
    std::vector<int32_t> gm;
    int64_t padBase = 0, vecBase = 0, colBase = 0, slotBase = 0, leafBase = 0;
    for (HMatrix* h : hs) {
        HMX_CHECK(h->b == (uint32_t)b, HMX_EINVAL, "all instances of a batch need the same bucketSize");
        HMX_CHECK(h->leafOffsets.size() == (size_t)h->S + 1, HMX_ESTATE, "H-matrix without leaf offsets");
        PlanHost H = plan_host(*h, false, chunk, 0, h->planDenseNear != 0);
        for (MvItem it : H.items) {
            it.xoff = (int32_t)(it.xoff + padBase);
            it.out = (int32_t)(it.out + slotBase);
            B.items.push_back(it);
        }
        for (int64_t p : H.colPart) B.colPart.push_back(p + slotBase);
        B.colNpart.insert(B.colNpart.end(), H.colNpart.begin(), H.colNpart.end());
        const int32_t e0 = (int32_t)B.entries.size();
        for (MvEntry en : H.entries) {
            en.zbase = (int32_t)(en.zbase + colBase);
            B.entries.push_back(en);
        }
        for (size_t q = 1; q < H.leafPtr.size(); ++q) B.leafPtr.push_back(e0 + H.leafPtr[q]);
        for (NearRef nr : H.nears) {
            nr.outLeaf += (int32_t)leafBase;
            nr.inLeaf += (int32_t)leafBase;
            B.nears.push_back(nr);
        }
        B.lowRankBytes += H.lowRankBytes;
        std::vector<int32_t> g = padding_map(*h, vecBase);
        gm.insert(gm.end(), g.begin(), g.end());
        padBase += (int64_t)h->S * b;
        vecBase += h->n;
        colBase += H.numCols;
        slotBase += H.numPart;
        leafBase += h->S;
    }
    HMX_CHECK(padBase < (int64_t)INT32_MAX && slotBase < (int64_t)INT32_MAX && colBase < (int64_t)INT32_MAX,
              HMX_EUNSUPPORTED, "batch too large for 32-bit plan indices");
    B.numCols = colBase;
    B.numPart = slotBase;
    upload_plan(c, P, B);
    P.chunk = chunk;
    P.n = vecBase;
    P.npad = padBase;
    P.gather = true;
    put(c, P.gmap, gm);
    diag.alloc((size_t)leafBase * b * b);
    size_t off = 0;
    for (HMatrix* h : hs) {
        HMX_HIP(hipMemcpyAsync(diag.p + off, h->diag.p, h->diag.bytes(), hipMemcpyDeviceToDevice, c.stream));
        off += h->diag.n;
    }
    P.diag = diag.p;
    HMX_HIP(hipStreamSynchronize(c.stream));
    P.built = true;
}

}  // namespace hmx
