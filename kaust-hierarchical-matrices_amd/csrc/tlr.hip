// TLR compression (createColumnMajorLRMatrix, src/constructTLRMatrix.cu:13-151)
// and Morton reordering (convertColumnMajorToMorton, src/helperFunctions.cu:8-57).
//
// One wave per b x b tile: the tile's points are gathered into LDS, the
// kernel tile exp(-|x-y|/ell) is evaluated once into LDS (no dense matrix in
// HBM), and the adaptive randomized approximation runs entirely in LDS:
//   Y = A*Omega (bs Gaussian samples from Philox), Y -= Q(Q^T Y) twice, then
//   column by column two modified Gram-Schmidt sweeps against the columns
//   already accepted from the block; a column is accepted when the norm of
//   that explicitly projected vector exceeds tau, and the ARA stops after r
//   consecutive small ones; accepted columns are projected against Q once
//   more, re-normalised and appended; finally V = A^T Q.
// The same algorithm, step for step, is the CPU oracle's ara_run().
#include <algorithm>

#include "hmx_device.h"
#include "hmx_internal.h"

namespace hmx {

constexpr int kBS = 16;  // ARA samples per block (kblas_ara_batch bs argument)

struct AraState {
    int k, small_run, converged, capped;
};

__device__ inline double wave_sum64(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;  // butterfly: every lane holds the identical sum
}

template <int B>
struct TileLds {
    static constexpr int MR = B;  // rank capacity
    double A[B * B];
    double Om[B * kBS];  // Omega, later the accepted block columns
    double Y[B * kBS];
    double Q[B * MR];
    double W[MR * kBS];
    double xr[3 * B], xc[3 * B];
    AraState st;
};

template <int B, int D>
__global__ void __launch_bounds__(64) k_tlr_ara(uint32_t n, uint32_t S, const double* __restrict__ p,
                                                const int32_t* __restrict__ perm, const int32_t* __restrict__ offsets,
                                                double ell, double tau, int r, int maxRank, uint64_t seed, uint64_t t0,
                                                double* __restrict__ Uslot, double* __restrict__ Vslot,
                                                int32_t* __restrict__ ranks, int32_t* __restrict__ caps,
                                                double* __restrict__ diag, const int32_t* __restrict__ tileMO) {
    extern __shared__ __align__(16) unsigned char smem[];
    TileLds<B>& L = *reinterpret_cast<TileLds<B>*>(smem);
    const int tid = threadIdx.x;
    // the tile: the next one of the column-major enumeration, or the next
    // entry of a Morton tile list; t (the column-major index u*S + v) is the
    // random stream id either way, so a tile draws the same Omega in both
    uint64_t t;
    uint32_t u, v;
    if (tileMO) {
        const uint32_t code = (uint32_t)tileMO[t0 + blockIdx.x];
        u = compact16(code);
        v = compact16(code >> 1);
        t = (uint64_t)u * S + v;
    } else {
        t = t0 + blockIdx.x;
        u = (uint32_t)(t / S);
        v = (uint32_t)(t % S);
    }
    const int au = offsets[u], su = offsets[u + 1] - au;
    const int av = offsets[v], sv = offsets[v + 1] - av;
    for (int e = tid; e < D * B; e += 64) {
        int j = e / B, c = e % B;
        L.xc[c * 3 + j] = c < su ? p[(size_t)j * n + perm[au + c]] : 0.0;
        L.xr[c * 3 + j] = c < sv ? p[(size_t)j * n + perm[av + c]] : 0.0;
    }
    __syncthreads();
    for (int e = tid; e < B * B; e += 64) {
        int c = e / B, q = e % B;
        double val;
        if (c < su && q < sv) val = kernel_eval<D>(&L.xc[c * 3], &L.xr[q * 3], ell);
        else val = (u == v && c == q) ? 1.0 : 0.0;
        L.A[e] = val;
    }
    __syncthreads();
    const size_t slot = (size_t)blockIdx.x;
    if (u == v) {  // dense diagonal block (generateDenseBlockColumn diagonal branch)
        if (diag)
            for (int e = tid; e < B * B; e += 64) diag[(size_t)u * B * B + e] = L.A[e];
        if (tid == 0) { ranks[slot] = 0; caps[slot] = 0; }
        return;
    }
    if (tid == 0) { L.st.k = 0; L.st.small_run = 0; L.st.converged = 0; L.st.capped = 0; }
    __syncthreads();
    // every block either accepts a column (k grows, bounded by maxRank) or
    // extends the small-column run (bounded by r), so this bound is never hit
    // by a finite tile; it only guards against NaN-poisoned input
    const int maxIters = (2 * maxRank + r) / kBS + 2;
    int it = 0;
    while (it < maxIters) {
        int k = L.st.k;
        if (L.st.converged || L.st.capped || k >= maxRank) break;
        // Omega
        for (int e = tid; e < B * kBS; e += 64) {
            int s = e / B, c = e % B;
            L.Om[e] = gauss_at(seed, kTagOmega, t, (uint32_t)(it * kBS + s), (uint32_t)c);
        }
        __syncthreads();
        // Y = A Omega
        for (int e = tid; e < B * kBS; e += 64) {
            int s = e / B, q = e % B;
            double acc = 0.0;
            for (int c = 0; c < B; ++c) acc += L.A[c * B + q] * L.Om[s * B + c];
            L.Y[e] = acc;
        }
        __syncthreads();
        for (int pass = 0; pass < 2 && k > 0; ++pass) {
            for (int e = tid; e < k * kBS; e += 64) {
                int s = e / k, a = e % k;
                double acc = 0.0;
                for (int q = 0; q < B; ++q) acc += L.Q[a * B + q] * L.Y[s * B + q];
                L.W[s * k + a] = acc;
            }
            __syncthreads();
            for (int e = tid; e < B * kBS; e += 64) {
                int s = e / B, q = e % B;
                double acc = L.Y[e];
                for (int a = 0; a < k; ++a) acc -= L.Q[a * B + q] * L.W[s * k + a];
                L.Y[e] = acc;
            }
            __syncthreads();
        }
        // column-by-column decisions on explicitly projected vectors; lane q owns row q
        double* Qb = L.Om;
        int na = 0, small_run = L.st.small_run, converged = 0, capped = 0;
        const int capleft = maxRank - k;
        const bool own = tid < B;
        for (int j = 0; j < kBS; ++j) {
            double y = own ? L.Y[j * B + tid] : 0.0;
            for (int pass = 0; pass < 2; ++pass)
                for (int a = 0; a < na; ++a) {
                    const double qa = own ? Qb[a * B + tid] : 0.0;
                    const double d = wave_sum64(qa * y);
                    y -= qa * d;
                }
            const double nu = sqrt(wave_sum64(y * y));
            if (nu > tau) {
                if (na >= capleft) { capped = 1; break; }
                if (own) Qb[na * B + tid] = y / nu;
                ++na;
                small_run = 0;
            } else if (++small_run >= r) {
                converged = 1;
                break;
            }
        }
        __syncthreads();
        if (na > 0) {
            if (k > 0) {  // accepted columns -= Q (Q^T accepted)
                for (int e = tid; e < k * na; e += 64) {
                    int s = e / k, a = e % k;
                    double acc = 0.0;
                    for (int q = 0; q < B; ++q) acc += L.Q[a * B + q] * Qb[s * B + q];
                    L.W[s * k + a] = acc;
                }
                __syncthreads();
                for (int e = tid; e < B * na; e += 64) {
                    int s = e / B, q = e % B;
                    double acc = Qb[e];
                    for (int a = 0; a < k; ++a) acc -= L.Q[a * B + q] * L.W[s * k + a];
                    Qb[e] = acc;
                }
                __syncthreads();
            }
            for (int j = 0; j < na; ++j) {  // re-normalise within the block
                double q = own ? Qb[j * B + tid] : 0.0;
                for (int a = 0; a < j; ++a) {
                    const double qa = own ? Qb[a * B + tid] : 0.0;
                    const double d = wave_sum64(qa * q);
                    q -= qa * d;
                }
                const double nu = sqrt(wave_sum64(q * q));
                if (own) {
                    q /= nu;
                    Qb[j * B + tid] = q;
                    L.Q[(k + j) * B + tid] = q;
                }
            }
        }
        __syncthreads();
        if (tid == 0) {
            L.st.k = k + na;
            L.st.small_run = small_run;
            L.st.converged = converged;
            L.st.capped = capped;
        }
        ++it;
        __syncthreads();
    }
    const int k = L.st.k;
    double* Us = Uslot + slot * (size_t)B * maxRank;
    double* Vs = Vslot + slot * (size_t)B * maxRank;
    for (int e = tid; e < k * B; e += 64) {
        int a = e / B, c = e % B;
        double acc = 0.0;
        for (int q = 0; q < B; ++q) acc += L.A[c * B + q] * L.Q[a * B + q];
        Vs[e] = acc;
        Us[e] = L.Q[e];
    }
    if (tid == 0) {
        ranks[slot] = k;
        caps[slot] = (L.st.capped || (!L.st.converged && k >= maxRank)) ? 1 : 0;
    }
}

// copy a batch's slots to packed storage: wave per tile
__global__ void __launch_bounds__(256) k_compact(uint64_t nt, int b, int slotCols, const int32_t* __restrict__ ranks,
                                                 const int64_t* __restrict__ off, const double* __restrict__ Us,
                                                 const double* __restrict__ Vs, double* __restrict__ U,
                                                 double* __restrict__ V) {
    uint64_t t = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= nt) return;
    int lane = threadIdx.x & 63;
    size_t cnt = (size_t)ranks[t] * b;
    const double* us = Us + t * (size_t)b * slotCols;
    const double* vs = Vs + t * (size_t)b * slotCols;
    double* ud = U + (size_t)off[t] * b;
    double* vd = V + (size_t)off[t] * b;
    for (size_t e = lane; e < cnt; e += 64) {
        ud[e] = us[e];
        vd[e] = vs[e];
    }
}

struct TlrLaunch {
    uint32_t n, S;
    const double* p;
    const int32_t* perm;
    const int32_t* off;
    double ell, tau;
    int r, maxRank;
    uint64_t seed;
    const int32_t* tileMO;  // Morton tile list, or null for the column-major enumeration
};

template <int B, int D>
static void launch_tlr(Context& c, const TlrLaunch& a, uint64_t t0, uint64_t nt, double* Us, double* Vs,
                       int32_t* ranks, int32_t* caps, double* diag) {
    size_t lds = sizeof(TileLds<B>);
    auto kern = k_tlr_ara<B, D>;
    if (lds > 64 * 1024) HMX_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    kern<<<(unsigned)nt, 64, lds, c.stream>>>(a.n, a.S, a.p, a.perm, a.off, a.ell, a.tau, a.r, a.maxRank, a.seed, t0,
                                              Us, Vs, ranks, caps, diag, a.tileMO);
    HMX_LAUNCHED();
}

template <int B>
static void launch_tlr_d(Context& c, uint32_t d, const TlrLaunch& a, uint64_t t0, uint64_t nt, double* Us, double* Vs,
                         int32_t* ranks, int32_t* caps, double* diag) {
    if (d == 1) launch_tlr<B, 1>(c, a, t0, nt, Us, Vs, ranks, caps, diag);
    else if (d == 2) launch_tlr<B, 2>(c, a, t0, nt, Us, Vs, ranks, caps, diag);
    else launch_tlr<B, 3>(c, a, t0, nt, Us, Vs, ranks, caps, diag);
}

static void launch_tlr_b(Context& c, uint32_t b, uint32_t d, const TlrLaunch& a, uint64_t t0, uint64_t nt, double* Us,
                         double* Vs, int32_t* ranks, int32_t* caps, double* diag) {
    switch (b) {
        case 8: launch_tlr_d<8>(c, d, a, t0, nt, Us, Vs, ranks, caps, diag); break;
        case 16: launch_tlr_d<16>(c, d, a, t0, nt, Us, Vs, ranks, caps, diag); break;
        case 32: launch_tlr_d<32>(c, d, a, t0, nt, Us, Vs, ranks, caps, diag); break;
        default: launch_tlr_d<64>(c, d, a, t0, nt, Us, Vs, ranks, caps, diag); break;
    }
}

double ara_threshold(double tol) { return tol / (10.0 * std::sqrt(2.0 / 3.141592653589793)); }

// Compress 'total' tiles in batches sized so the per-tile slots stay under
// ~2 GiB; ranks[total] receives the ranks, U/V the packed factors in tile
// order.  Returns the number of capped tiles.
static uint64_t compress_tiles(Context& c, const hmx_config& cfg, const KDTree& t, const TlrLaunch& a,
                               uint64_t total, int32_t* ranks, double* diag, DBuf<double>& U, DBuf<double>& V,
                               uint64_t& rankSum) {
    const uint32_t b = t.b;
    const int maxRank = a.maxRank;
    const size_t slotBytes = 2 * (size_t)b * maxRank * sizeof(double);
    uint64_t batch = std::max<uint64_t>(1, ((size_t)2 << 30) / std::max<size_t>(slotBytes, 1));
    batch = std::min<uint64_t>(batch, std::max<uint64_t>(total, 1));
    DBuf<double> Us(batch * b * maxRank), Vs(batch * b * maxRank);
    DBuf<int32_t> caps(batch);
    DBuf<int64_t> boff(batch + 1);
    std::vector<DBuf<double>> partsU, partsV;
    std::vector<uint64_t> partBase;
    std::vector<int32_t> hcaps(batch);
    uint64_t running = 0, capHits = 0;
    for (uint64_t t0 = 0; t0 < total; t0 += batch) {
        const uint64_t nt = std::min<uint64_t>(batch, total - t0);
        int32_t* rk = ranks + t0;
        launch_tlr_b(c, b, t.d, a, t0, nt, Us.p, Vs.p, rk, caps.p, diag);
        exclusive_scan_i32_i64(c, rk, boff.p, nt);
        int64_t btotal = 0;
        HMX_HIP(hipMemcpyAsync(&btotal, boff.p + nt, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
        HMX_HIP(hipMemcpyAsync(hcaps.data(), caps.p, nt * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
        HMX_HIP(hipStreamSynchronize(c.stream));
        for (uint64_t i = 0; i < nt; ++i) capHits += (uint64_t)hcaps[i];
        partsU.emplace_back((size_t)btotal * b);
        partsV.emplace_back((size_t)btotal * b);
        partBase.push_back(running);
        if (btotal > 0) {
            k_compact<<<(unsigned)((nt + 3) / 4), 256, 0, c.stream>>>(nt, (int)b, maxRank, rk, boff.p, Us.p, Vs.p,
                                                                      partsU.back().p, partsV.back().p);
            HMX_LAUNCHED();
        }
        running += (uint64_t)btotal;
        HMX_HIP(hipStreamSynchronize(c.stream));
    }
    Us.release();
    Vs.release();
    rankSum = running;
    U.alloc((size_t)running * b);
    V.alloc((size_t)running * b);
    for (size_t i = 0; i < partsU.size(); ++i) {
        if (!partsU[i].n) continue;
        HMX_HIP(hipMemcpyAsync(U.p + partBase[i] * b, partsU[i].p, partsU[i].bytes(), hipMemcpyDeviceToDevice,
                               c.stream));
        HMX_HIP(hipMemcpyAsync(V.p + partBase[i] * b, partsV[i].p, partsV[i].bytes(), hipMemcpyDeviceToDevice,
                               c.stream));
    }
    HMX_HIP(hipStreamSynchronize(c.stream));
    return capHits;
}

static TlrLaunch tlr_args(const hmx_config& cfg, const KDTree& t, const double* d_points, const int32_t* tileMO) {
    TlrLaunch a;
    a.n = t.n;
    a.S = t.numSegments;
    a.p = d_points;
    a.perm = t.perm.p;
    a.off = t.offsets.p;
    a.ell = cfg.correlationLength > 0 ? cfg.correlationLength : default_ell(t.d);
    a.tau = ara_threshold((double)cfg.lowestLevelTolerance);
    a.r = cfg.araR;
    a.maxRank = cfg.maxRankPolicy == HMX_MAXRANK_REF ? (int)t.b / 2 : (int)t.b;
    a.seed = cfg.seed;
    a.tileMO = tileMO;
    return a;
}

static void check_tlr_args(const hmx_config& cfg, const KDTree& t) {
    const uint32_t b = t.b;
    HMX_CHECK(b == 8 || b == 16 || b == 32 || b == 64, HMX_EUNSUPPORTED,
              "TLR tiles: bucketSize must be 8, 16, 32 or 64");
    HMX_CHECK(cfg.araBlockSize == kBS, HMX_EUNSUPPORTED, "ARA block size must be 16");
    HMX_CHECK(cfg.araR >= 1, HMX_EINVAL, "araR must be >= 1");
    HMX_CHECK(t.numSegments >= 2, HMX_EINVAL, "TLR needs at least two segments (n > bucketSize)");
}

void build_tlr_cm(Context& c, const hmx_config& cfg, const KDTree& t, const double* d_points, TLR& out) {
    check_tlr_args(cfg, t);
    const uint32_t b = t.b, S = t.numSegments;
    out.ordering = HMX_COLUMN_MAJOR;
    out.S = S;
    out.b = b;
    out.n = t.n;
    const uint64_t total = (uint64_t)S * S;
    out.ranks.alloc(total);
    out.offsets.alloc(total + 1);
    out.diagonal.alloc((size_t)S * b * b);
    const TlrLaunch a = tlr_args(cfg, t, d_points, nullptr);
    uint64_t rankSum = 0;
    out.capHits = compress_tiles(c, cfg, t, a, total, out.ranks.p, out.diagonal.p, out.U, out.V, rankSum);
    out.rankSum = rankSum;
    exclusive_scan_i32_i64(c, out.ranks.p, out.offsets.p, total);
    HMX_HIP(hipStreamSynchronize(c.stream));
}

// Leaf-level tiles of a structure compressed straight from the kernel (the
// TLR ARA over a Morton tile list): the same factors, bit for bit, as the
// TLR matrix's tiles at those positions.
void compress_leaf_tiles(Context& c, const hmx_config& cfg, const KDTree& t, const double* d_points,
                         const std::vector<int32_t>& tiles, HLevel& out) {
    check_tlr_args(cfg, t);
    const int T = (int)tiles.size();
    out.numTiles = T;
    out.m = (int)t.b;
    out.h_ranks.assign(T, 0);
    out.h_offsets.assign(T + 1, 0);
    out.rankSum = 0;
    out.capHits = 0;
    if (T == 0) return;
    DBuf<int32_t> d_tiles(T), d_r(T);
    HMX_HIP(hipMemcpyAsync(d_tiles.p, tiles.data(), T * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
    const TlrLaunch a = tlr_args(cfg, t, d_points, d_tiles.p);
    uint64_t rankSum = 0;
    out.capHits = (int32_t)compress_tiles(c, cfg, t, a, (uint64_t)T, d_r.p, nullptr, out.U, out.V, rankSum);
    HMX_HIP(hipMemcpyAsync(out.h_ranks.data(), d_r.p, T * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    HMX_HIP(hipStreamSynchronize(c.stream));
    int64_t acc = 0;
    for (int q = 0; q < T; ++q) { out.h_offsets[q] = acc; acc += out.h_ranks[q]; }
    out.h_offsets[T] = acc;
    out.rankSum = acc;
    HMX_CHECK((uint64_t)acc == rankSum, HMX_EDEVICE, "leaf tile compression: rank sum mismatch");
}

// Dense diagonal blocks, exactly as the TLR kernel writes them: D[leaf] is
// column-major b x b, K(x_c, x_q) for real points, identity on the padding.
template <int D>
__global__ void __launch_bounds__(256) k_dense_diag(uint32_t n, int b, uint32_t S, const double* __restrict__ p,
                                                    const int32_t* __restrict__ perm,
                                                    const int32_t* __restrict__ offsets, double ell,
                                                    double* __restrict__ diag) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t bb = (int64_t)b * b;
    if (e >= (int64_t)S * bb) return;
    const int leaf = (int)(e / bb);
    const int c = (int)((e % bb) / b), q = (int)(e % b);
    const int a0 = offsets[leaf], cnt = offsets[leaf + 1] - a0;
    double val;
    if (c < cnt && q < cnt) {
        double xc[3], xq[3];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            xc[j] = p[(size_t)j * n + perm[a0 + c]];
            xq[j] = p[(size_t)j * n + perm[a0 + q]];
        }
        val = kernel_eval<D>(xc, xq, ell);
    } else {
        val = c == q ? 1.0 : 0.0;
    }
    diag[e] = val;
}

void build_dense_diagonal(Context& c, const hmx_config& cfg, const KDTree& t, const double* d_points, double* diag) {
    const double ell = cfg.correlationLength > 0 ? cfg.correlationLength : default_ell(t.d);
    const int64_t tot = (int64_t)t.numSegments * t.b * t.b;
    const unsigned g = (unsigned)((tot + 255) / 256);
    if (t.d == 1) k_dense_diag<1><<<g, 256, 0, c.stream>>>(t.n, (int)t.b, t.numSegments, d_points, t.perm.p, t.offsets.p, ell, diag);
    else if (t.d == 2) k_dense_diag<2><<<g, 256, 0, c.stream>>>(t.n, (int)t.b, t.numSegments, d_points, t.perm.p, t.offsets.p, ell, diag);
    else k_dense_diag<3><<<g, 256, 0, c.stream>>>(t.n, (int)t.b, t.numSegments, d_points, t.perm.p, t.offsets.p, ell, diag);
    HMX_LAUNCHED();
}

// ---- column-major -> Morton ------------------------------------------------
__global__ void __launch_bounds__(256) k_ranks_to_mo(uint32_t S, const int32_t* __restrict__ cm,
                                                     int32_t* __restrict__ mo) {
    uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (uint64_t)S * S) return;
    mo[cm_to_mo(S, (uint32_t)i)] = cm[i];
}

__global__ void __launch_bounds__(256) k_tiles_to_mo(uint32_t S, int b, const int32_t* __restrict__ ranks,
                                                     const int64_t* __restrict__ offCM,
                                                     const int64_t* __restrict__ offMO, const double* __restrict__ Ucm,
                                                     const double* __restrict__ Vcm, double* __restrict__ Umo,
                                                     double* __restrict__ Vmo) {
    uint64_t i = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= (uint64_t)S * S) return;
    int lane = threadIdx.x & 63;
    size_t cnt = (size_t)ranks[i] * b;
    uint32_t mo = cm_to_mo(S, (uint32_t)i);
    const double* us = Ucm + (size_t)offCM[i] * b;
    const double* vs = Vcm + (size_t)offCM[i] * b;
    double* ud = Umo + (size_t)offMO[mo] * b;
    double* vd = Vmo + (size_t)offMO[mo] * b;
    for (size_t e = lane; e < cnt; e += 64) {
        ud[e] = us[e];
        vd[e] = vs[e];
    }
}

void convert_to_morton(Context& c, const TLR& cm, TLR& mo) {
    HMX_CHECK(cm.ordering == HMX_COLUMN_MAJOR, HMX_ESTATE, "convertColumnMajorToMorton: input is not column-major");
    HMX_CHECK(cm.S <= 65536, HMX_EINVAL, "Morton codes are 32-bit: numSegments must be <= 65536");
    const uint32_t S = cm.S, b = cm.b;
    const uint64_t total = (uint64_t)S * S;
    mo.ordering = HMX_MORTON;
    mo.S = S;
    mo.b = b;
    mo.n = cm.n;
    mo.rankSum = cm.rankSum;
    mo.capHits = cm.capHits;
    mo.ranks.alloc(total);
    mo.offsets.alloc(total + 1);
    mo.U.alloc(cm.U.n);
    mo.V.alloc(cm.V.n);
    mo.diagonal.alloc(cm.diagonal.n);
    k_ranks_to_mo<<<(unsigned)((total + 255) / 256), 256, 0, c.stream>>>(S, cm.ranks.p, mo.ranks.p);
    HMX_LAUNCHED();
    exclusive_scan_i32_i64(c, mo.ranks.p, mo.offsets.p, total);
    k_tiles_to_mo<<<(unsigned)((total + 3) / 4), 256, 0, c.stream>>>(S, (int)b, cm.ranks.p, cm.offsets.p,
                                                                    mo.offsets.p, cm.U.p, cm.V.p, mo.U.p, mo.V.p);
    HMX_LAUNCHED();
    HMX_HIP(hipMemcpyAsync(mo.diagonal.p, cm.diagonal.p, cm.diagonal.bytes(), hipMemcpyDeviceToDevice, c.stream));
    HMX_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace hmx
