// H-matrix x vector (HMatrixVecMult, src/HMatrixVectorMultiplication.cu:10-26;
// cutlassDiagonalXVec, src/cutlassDiagonalXVector.cu:54-76; cutlassHierarchicalXVec,
// src/cutlassHMatrixXVector.cu:330-398).
//
// y = D x + sum over all levels and tiles t of U_t (V_t^T x_col(t)), for any
// admissibility structure, in three launches (all levels at once):
//   phase 1  z_t = V_t^T x : one wave per work item, a contiguous run of V
//            read with 16-byte loads; long columns are split into 2048-element
//            chunks that produce partial sums, short columns are packed many
//            per item and reduced with segmented wave reductions;
//   reduce   z = sum of the chunk partials (fixed order -> deterministic);
//   phase 2  one workgroup per leaf (32 or 64 rows): y_leaf = D_leaf x_leaf + sum of the
//            leaf's rows of every U_t covering it times z_t; each y element is
//            written exactly once, no atomics.
// The factors stay where the build left them (no per-call V transpose as in
// src/cutlassHMatrixXVector.cu:118-150, no per-call host problem assembly).
#include <algorithm>

#include "hmx_device.h"
#include "hmx_internal.h"

namespace hmx {

constexpr int kChunk = 2048;  // elements of one single-column work item

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int N>
__device__ inline void vload_d(const double* __restrict__ p, double (&out)[N]) {
#pragma unroll
    for (int i = 0; i < N; i += 2) {
        const double2 v = *reinterpret_cast<const double2*>(p + i);
        out[i] = v.x;
        out[i + 1] = v.y;
    }
}

// ---- phase 1 ---------------------------------------------------------------
template <int K>
__global__ void __launch_bounds__(256) k_mv_phase1(const MvItem* __restrict__ items, int numItems,
                                                   const double* __restrict__ X, int ldx,
                                                   double* __restrict__ zpart) {
    const int lane = threadIdx.x & 63;
    const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wid >= numItems) return;
    const MvItem it = items[wid];
    const int len = it.len;
    const int m = it.colLen ? it.colLen : len;  // column length inside this item (power of two)
    const double* __restrict__ xs = X + it.xoff;
    double acc[K];
#pragma unroll
    for (int s = 0; s < K; ++s) acc[s] = 0.0;
    if (m >= 128) {
        // every 128-element step lies inside one column
        int col = 0;
        for (int e0 = 0; e0 < len; e0 += 128) {
            const int e = e0 + 2 * lane;
            const double2 vv = *reinterpret_cast<const double2*>(it.v + e);
            const int xi = e & (m - 1);
#pragma unroll
            for (int s = 0; s < K; ++s) {
                const double2 xx = *reinterpret_cast<const double2*>(xs + (size_t)s * ldx + xi);
                acc[s] = fma(vv.x, xx.x, acc[s]);
                acc[s] = fma(vv.y, xx.y, acc[s]);
            }
            if (((e0 + 128) & (m - 1)) == 0) {
#pragma unroll
                for (int s = 0; s < K; ++s) {
                    double r = wave_sum(acc[s]);
                    if (lane == 0) zpart[(size_t)(it.out + col) * K + s] = r;
                    acc[s] = 0.0;
                }
                ++col;
            }
        }
    } else {
        // m in {32, 64}: one step covers 128/m columns; m/2 lanes per column
        const int g = m >> 1;
        for (int e0 = 0; e0 < len; e0 += 128) {
            const int e = e0 + 2 * lane;
            const bool ok = e < len;
            double2 vv = make_double2(0.0, 0.0);
            if (ok) vv = *reinterpret_cast<const double2*>(it.v + e);
            const int xi = e & (m - 1);
            double p[K];
#pragma unroll
            for (int s = 0; s < K; ++s) {
                double2 xx = make_double2(0.0, 0.0);
                if (ok) xx = *reinterpret_cast<const double2*>(xs + (size_t)s * ldx + xi);
                p[s] = fma(vv.y, xx.y, vv.x * xx.x);
            }
            for (int off = 1; off < g; off <<= 1)
#pragma unroll
                for (int s = 0; s < K; ++s) p[s] += __shfl_xor(p[s], off, 64);
            if (ok && (lane & (g - 1)) == 0) {
                const int col = e / m;
#pragma unroll
                for (int s = 0; s < K; ++s) zpart[(size_t)(it.out + col) * K + s] = p[s];
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
                                                   const double* __restrict__ X, int ldx, double* __restrict__ Y) {
    constexpr int RPL = BL / 16;
    static_assert(RPL == 2 || RPL == 4, "bucketSize 32 or 64");
    __shared__ double red[4][BL * K];
    const int leaf = blockIdx.x;
    const int l16 = threadIdx.x & 15, cg = threadIdx.x >> 4;
    const int r0 = RPL * l16;
    double acc[RPL][K];
#pragma unroll
    for (int q = 0; q < RPL; ++q)
#pragma unroll
        for (int s = 0; s < K; ++s) acc[q][s] = 0.0;
    // dense diagonal block (column-major): columns c = cg, cg+16, ...
    const double* D = diag + (size_t)leaf * BL * BL;
    const double* xl = X + (size_t)leaf * BL;
#pragma unroll
    for (int cc = 0; cc < BL / 16; ++cc) {
        const int c = cg + 16 * cc;
        double dv[RPL];
        vload_d<RPL>(D + c * BL + r0, dv);
#pragma unroll
        for (int s = 0; s < K; ++s) {
            const double xv = xl[(size_t)s * ldx + c];
#pragma unroll
            for (int q = 0; q < RPL; ++q) acc[q][s] = fma(dv[q], xv, acc[q][s]);
        }
    }
    const int e0 = leafPtr[leaf], e1 = leafPtr[leaf + 1];
    for (int e = e0; e < e1; ++e) {
        const MvEntry en = entries[e];
        const double* __restrict__ U = en.u;
        const double* __restrict__ zz = z + (size_t)en.zbase * K;
        const int m = en.m;
        int j = cg;
        for (; j + 48 < en.rank; j += 64) {
            double u[4][RPL];
#pragma unroll
            for (int p = 0; p < 4; ++p) vload_d<RPL>(U + (size_t)(j + 16 * p) * m + r0, u[p]);
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int s = 0; s < K; ++s) {
                    const double zv = zz[(size_t)(j + 16 * p) * K + s];
#pragma unroll
                    for (int q = 0; q < RPL; ++q) acc[q][s] = fma(u[p][q], zv, acc[q][s]);
                }
        }
        for (; j < en.rank; j += 16) {
            double u[RPL];
            vload_d<RPL>(U + (size_t)j * m + r0, u);
#pragma unroll
            for (int s = 0; s < K; ++s) {
                const double zv = zz[(size_t)j * K + s];
#pragma unroll
                for (int q = 0; q < RPL; ++q) acc[q][s] = fma(u[q], zv, acc[q][s]);
            }
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
            for (int q = 0; q < RPL; ++q) red[wave][s * BL + r0 + q] = acc[q][s];
    __syncthreads();
    for (int o = threadIdx.x; o < BL * K; o += 256) {
        const double v = (red[0][o] + red[1][o]) + (red[2][o] + red[3][o]);
        const int s = o / BL, row = o % BL;
        Y[(size_t)s * ldx + (size_t)leaf * BL + row] = v;
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

void build_matvec_plan(Context& c, HMatrix& h, bool transpose) {
    MatvecPlan& P = transpose ? h.planT : h.plan;
    const int L = h.numLevels, b = (int)h.b, S = (int)h.S;
    HMX_CHECK(h.n == (uint32_t)S * (uint32_t)b, HMX_EUNSUPPORTED,
              "HMatrixVecMult requires full leaves (n = numSegments * bucketSize), as the reference does");
    HMX_CHECK(b == 32 || b == 64, HMX_EUNSUPPORTED, "HMatrixVecMult kernels are built for bucketSize 32 or 64");
    std::vector<MvItem> items;
    std::vector<int64_t> colPart;
    std::vector<int32_t> colNpart;
    std::vector<std::vector<MvEntry>> leafEntries(S);
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
            const int nch = m >= kChunk ? m / kChunk : 1;
            for (int j = 0; j < r; ++j) {
                colPart.push_back(slot + (int64_t)j * nch);
                colNpart.push_back(nch);
            }
            if (m >= kChunk) {
                for (int j = 0; j < r; ++j)
                    for (int q = 0; q < nch; ++q)
                        items.push_back(MvItem{Pin + (size_t)j * m + (size_t)q * kChunk, (int32_t)(inCl * m + q * kChunk),
                                               kChunk, 0, (int32_t)(slot + (int64_t)j * nch + q)});
            } else {
                const int cpi = kChunk / m;
                for (int j0 = 0; j0 < r; j0 += cpi) {
                    const int nc = std::min(cpi, r - j0);
                    items.push_back(MvItem{Pin + (size_t)j0 * m, (int32_t)(inCl * m), nc * m, m, (int32_t)(slot + j0)});
                }
            }
            for (int q = 0; q < bus; ++q)
                leafEntries[outCl * bus + q].push_back(MvEntry{Pout + (size_t)q * b, m, r, (int32_t)g, 0});
            slot += (int64_t)r * nch;
            g += r;
        }
    }
    HMX_CHECK(slot < (int64_t)INT32_MAX && g < (int64_t)INT32_MAX, HMX_EUNSUPPORTED, "matvec plan too large");
    std::vector<int32_t> leafPtr(S + 1, 0);
    std::vector<MvEntry> entries;
    for (int leaf = 0; leaf < S; ++leaf) {
        leafPtr[leaf] = (int32_t)entries.size();
        entries.insert(entries.end(), leafEntries[leaf].begin(), leafEntries[leaf].end());
    }
    leafPtr[S] = (int32_t)entries.size();
    P.numItems = (int32_t)items.size();
    P.numCols = g;
    P.numPart = slot;
    P.items.alloc(std::max<size_t>(1, items.size()));
    P.colPart.alloc(std::max<size_t>(1, colPart.size()));
    P.colNpart.alloc(std::max<size_t>(1, colNpart.size()));
    P.leafPtr.alloc(S + 1);
    P.entries.alloc(std::max<size_t>(1, entries.size()));
    HMX_HIP(hipMemcpyAsync(P.items.p, items.data(), items.size() * sizeof(MvItem), hipMemcpyHostToDevice, c.stream));
    HMX_HIP(hipMemcpyAsync(P.colPart.p, colPart.data(), colPart.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                           c.stream));
    HMX_HIP(hipMemcpyAsync(P.colNpart.p, colNpart.data(), colNpart.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                           c.stream));
    HMX_HIP(hipMemcpyAsync(P.leafPtr.p, leafPtr.data(), leafPtr.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                           c.stream));
    HMX_HIP(hipMemcpyAsync(P.entries.p, entries.data(), entries.size() * sizeof(MvEntry), hipMemcpyHostToDevice,
                           c.stream));
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
static void run_phases(Context& c, MatvecPlan& P, int S, int b, const double* X, int ldx, double* Y,
                       hipEvent_t* ev) {
    if (ev) HMX_HIP(hipEventRecord(ev[0], c.stream));
    if (P.numItems > 0) {
        k_mv_phase1<K><<<(unsigned)((P.numItems + 3) / 4), 256, 0, c.stream>>>(P.items.p, P.numItems, X, ldx,
                                                                                 P.zpart.p);
        HMX_LAUNCHED();
    }
    if (ev) HMX_HIP(hipEventRecord(ev[1], c.stream));
    if (P.numCols > 0) {
        k_mv_reduce<<<(unsigned)((P.numCols * K + 255) / 256), 256, 0, c.stream>>>(P.numCols, K, P.colPart.p,
                                                                                   P.colNpart.p, P.zpart.p, P.z.p);
        HMX_LAUNCHED();
    }
    if (ev) HMX_HIP(hipEventRecord(ev[2], c.stream));
    if (b == 64)
        k_mv_phase2<64, K><<<(unsigned)S, 256, 0, c.stream>>>(P.leafPtr.p, P.entries.p, P.diag, P.z.p, X, ldx, Y);
    else
        k_mv_phase2<32, K><<<(unsigned)S, 256, 0, c.stream>>>(P.leafPtr.p, P.entries.p, P.diag, P.z.p, X, ldx, Y);
    HMX_LAUNCHED();
    if (ev) HMX_HIP(hipEventRecord(ev[3], c.stream));
}

template <int K>
static void run_k(Context& c, MatvecPlan& P, HMatrix& h, uint32_t k, const double* X, double* Y, hipEvent_t* ev) {
    const int ldx = (int)h.n;
    for (uint32_t s0 = 0; s0 < k; s0 += K)
        run_phases<K>(c, P, (int)h.S, (int)h.b, X + (size_t)s0 * ldx, ldx, Y + (size_t)s0 * ldx, ev);
}

void hmatvec(Context& c, HMatrix& h, uint32_t k, const double* d_in, double* d_out, bool transpose, float* ms,
             int iters) {
    HMX_CHECK(k >= 1, HMX_EINVAL, "vectorWidth must be >= 1");
    HMX_CHECK(d_in && d_out, HMX_EINVAL, "null vector pointer");
    HMX_CHECK(d_in != d_out, HMX_EINVAL, "in and out must not alias");
    MatvecPlan& P = transpose ? h.planT : h.plan;
    if (!P.built) build_matvec_plan(c, h, transpose);
    // kernel width: 1, 4, 8 or 16 right-hand sides per pass
    const uint32_t K = k == 1 ? 1 : (k <= 4 ? 4 : (k <= 8 ? 8 : 16));
    HMX_CHECK(k == 1 || k % K == 0, HMX_EINVAL, "vectorWidth must be 1 or a multiple of 4/8/16");
    if (P.wsK < K) {
        P.zpart.alloc((size_t)std::max<int64_t>(1, P.numPart) * K);
        P.z.alloc((size_t)std::max<int64_t>(1, P.numCols) * K);
        P.wsK = K;
    }
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    const bool prof = ms != nullptr;
    if (prof)
        for (auto& e : ev) HMX_HIP(hipEventCreate(&e));
    const int reps = prof ? std::max(1, iters) : 1;
    float acc[4] = {0, 0, 0, 0};
    for (int it = 0; it < reps; ++it) {
        switch (K) {
            case 1: run_k<1>(c, P, h, k, d_in, d_out, prof ? ev : nullptr); break;
            case 4: run_k<4>(c, P, h, k, d_in, d_out, prof ? ev : nullptr); break;
            case 8: run_k<8>(c, P, h, k, d_in, d_out, prof ? ev : nullptr); break;
            default: run_k<16>(c, P, h, k, d_in, d_out, prof ? ev : nullptr); break;
        }
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

}  // namespace hmx
