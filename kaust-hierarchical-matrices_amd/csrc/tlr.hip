// TLR compression (createColumnMajorLRMatrix, src/constructTLRMatrix.cu:13-151)
// and Morton reordering (convertColumnMajorToMorton, src/helperFunctions.cu:8-57).
//
// One wave per b x b tile.  The tile's points are staged in LDS and the kernel
// tile A = exp(-|x-y|/ell) is never stored: every product with A evaluates its
// entries straight into the matrix-core operand registers.  The adaptive
// randomized approximation:
//   Y = A*Omega (bs Gaussian samples from Philox), Y -= Q(Q^T Y) twice, then
//   column by column two modified Gram-Schmidt sweeps against the columns
//   already accepted from the block; a column is accepted when the norm of
//   that explicitly projected vector exceeds tau, and the ARA stops after r
//   consecutive small ones; accepted columns are projected against Q once
//   more, re-normalised and appended; finally V = A^T Q.
// The block products (A*Omega, Q^T Y, Q W, A^T Q) run on the FP64 matrix
// cores (v_mfma_f64_16x16x4f64): the f64 output layout (lane l, register v
// holds row 4v + l/16, column l%16 of a 16 x 16 block) is exactly the B
// operand layout of the next product (lane l holds row 4ks + l/16, column
// l%16 of step ks), so Y and W stay in registers between products.  Q lives
// in LDS with an odd leading dimension, so the transposed operand reads of
// 16 lanes land on 16 different banks.  The column decisions run one row per
// lane with DPP wave sums.  The same algorithm, step for step, is the CPU
// oracle's ara_run().
#include <algorithm>

#include "hmx_device.h"
#include "hmx_internal.h"
#include "mfma.h"

namespace hmx {

constexpr int kBS = 16;  // ARA samples per block (kblas_ara_batch bs argument)

template <int B>
struct TileLds {
    static constexpr int BP = B < 16 ? 16 : B;  // rows padded to whole 16-row operand blocks
    static constexpr int LQ = BP + 1;           // odd leading dimension
    double Q[BP * LQ];                          // accepted basis, column a at Q[a * LQ]
    double S[kBS * LQ];                         // one block's columns (Y, then the accepted block)
    double xc[BP * 3], xr[BP * 3];              // column cluster (u) and row cluster (v) points
};

// Per-cloud constants of a TLR launch
struct TlrCloud {
    int64_t leafBase;  // first leaf of the cloud in the padded leaf space
    int64_t first;     // diagonal launches: first leaf of the cloud among the launch's leaves
    uint32_t S;        // leaves of the cloud (random stream ids are u * S + v)
    int maxRank;
    double ell, tau;
};

// Points of a launch: either one cloud's SoA coordinates p[j*n + perm[i]]
// with leaf offsets (the TLR matrix), or the padded leaf space P of one or
// more clouds, point i of leaf k of a cloud at P[(leafBase + k) * b + i] as
// (x, y, z, w) with w = 0 on the padding (direct and batched assembly)
struct TlrArgs {
    uint32_t n;
    const double* p;
    const int32_t* perm;
    const int32_t* offsets;
    const double4* P;
    const TlrCloud* clouds;     // per-cloud constants
    const int32_t* tileCloud;   // cloud of each listed tile, or null for cloud 0
    const int32_t* tileMO;      // Morton tile list (cloud-local codes), or null for column-major enumeration
    uint32_t diagLeaf0;         // diagonal block of leaf u goes to diag[(u - diagLeaf0) * b * b]
    int r;
    uint64_t seed;
};

template <int B, int D>
__global__ void __launch_bounds__(64) k_tlr_ara(TlrArgs g, uint64_t t0, double* __restrict__ Uslot,
                                                double* __restrict__ Vslot, int slotCols, int32_t* __restrict__ ranks,
                                                int32_t* __restrict__ caps, double* __restrict__ diag) {
    using L_t = TileLds<B>;
    constexpr int BP = L_t::BP, LQ = L_t::LQ, RB = BP / 16, KS = BP / 4;
    extern __shared__ __align__(16) unsigned char smem[];
    L_t& L = *reinterpret_cast<L_t*>(smem);
    const int tid = threadIdx.x, lr = tid & 15, lg = tid >> 4;
    const TlrCloud cl = g.clouds[g.tileCloud ? g.tileCloud[t0 + blockIdx.x] : 0];
    const uint32_t S = cl.S;
    const double ell = cl.ell, tau = cl.tau;
    const int maxRank = cl.maxRank, r = g.r;
    const uint64_t seed = g.seed;
    const double* __restrict__ p = g.p;
    const int32_t* __restrict__ perm = g.perm;
    const uint32_t n = g.n;
    // the tile: the next one of the column-major enumeration, or the next
    // entry of a Morton tile list; t (the column-major index u*S + v) is the
    // random stream id either way, so a tile draws the same Omega in both
    uint64_t t;
    uint32_t u, v;
    if (g.tileMO) {
        const uint32_t code = (uint32_t)g.tileMO[t0 + blockIdx.x];
        u = compact16(code);
        v = compact16(code >> 1);
        t = (uint64_t)u * S + v;
    } else {
        t = t0 + blockIdx.x;
        u = (uint32_t)(t / S);
        v = (uint32_t)(t % S);
    }
    int au = 0, su, av = 0, sv;
    if (g.P) {  // the first su positions of a padded leaf are its points
        const double4* Pu = g.P + (cl.leafBase + u) * B;
        const double4* Pv = g.P + (cl.leafBase + v) * B;
        su = __popcll(__ballot(tid < B && Pu[tid < B ? tid : 0].w != 0.0));
        sv = __popcll(__ballot(tid < B && Pv[tid < B ? tid : 0].w != 0.0));
    } else {
        au = g.offsets[u];
        su = g.offsets[u + 1] - au;
        av = g.offsets[v];
        sv = g.offsets[v + 1] - av;
    }
    const size_t slot = (size_t)blockIdx.x;
    if (u == v && !g.P) {  // dense diagonal block (generateDenseBlockColumn diagonal branch)
        if (diag)
            for (int e = tid; e < B * B; e += 64) {
                const int c = e / B, q = e % B;
                double val;
                if (c < su && q < sv) {
                    double xc[3], xq[3];
#pragma unroll
                    for (int j = 0; j < D; ++j) {
                        xc[j] = p[(size_t)j * n + perm[au + c]];
                        xq[j] = p[(size_t)j * n + perm[av + q]];
                    }
                    val = kernel_eval<D>(xc, xq, ell);
                } else {
                    val = c == q ? 1.0 : 0.0;
                }
                diag[(size_t)(u - g.diagLeaf0) * B * B + e] = val;
            }
        if (tid == 0) { ranks[slot] = 0; caps[slot] = 0; }
        return;
    }
    if (u == v) {  // listed diagonal position: nothing to compress
        if (tid == 0) { ranks[slot] = 0; caps[slot] = 0; }
        return;
    }
    if (g.P) {
        for (int e = tid; e < D * BP; e += 64) {
            const int j = e / BP, c = e % BP;
            const double* pu = reinterpret_cast<const double*>(g.P + (cl.leafBase + u) * B + c);
            const double* pv = reinterpret_cast<const double*>(g.P + (cl.leafBase + v) * B + c);
            L.xc[c * 3 + j] = c < su ? pu[j] : 0.0;
            L.xr[c * 3 + j] = c < sv ? pv[j] : 0.0;
        }
    } else {
        for (int e = tid; e < D * BP; e += 64) {
            const int j = e / BP, c = e % BP;
            L.xc[c * 3 + j] = c < su ? p[(size_t)j * n + perm[au + c]] : 0.0;
            L.xr[c * 3 + j] = c < sv ? p[(size_t)j * n + perm[av + c]] : 0.0;
        }
    }
    for (int e = tid; e < BP * LQ; e += 64) L.Q[e] = 0.0;
    __syncthreads();
    const double nil = -1.0 / ell;
    // the points of this lane's A-operand rows (row cluster v): rows 16 rb + lr
    double xrow[RB][D];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int j = 0; j < D; ++j) xrow[rb][j] = L.xr[(16 * rb + lr) * 3 + j];
    // A[q][c] = K(x_c, x_q) for real points, 0 on the padding
    auto entry = [&](const double* xq, int q, int c) -> double {
        return (c < su && q < sv) ? kernel_eval_scaled<D>(&L.xc[c * 3], xq, nil) : 0.0;
    };

    int k = 0, small_run = 0, converged = 0, capped = 0;
    // every block either accepts a column (k grows, bounded by maxRank) or
    // extends the small-column run (bounded by r), so this bound is never hit
    // by a finite tile; it only guards against NaN-poisoned input
    const int maxIters = (2 * maxRank + r) / kBS + 2;
    const bool own = tid < B;
    for (int it = 0; it < maxIters && !converged && !capped && k < maxRank; ++it) {
        // Y = A Omega: lane l holds Omega[4 ks + lg][lr] for step ks
        v4d y[RB];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) y[rb] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
        for (int ks = 0; ks < KS; ++ks) {
            const int c = 4 * ks + lg;
            const double om = c < B ? gauss_at(seed, kTagOmega, t, (uint32_t)(it * kBS + lr), (uint32_t)c) : 0.0;
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
                const double a = entry(xrow[rb], 16 * rb + lr, c);
                y[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, om, y[rb], 0, 0, 0);
            }
        }
        // Y -= Q (Q^T Y), twice; columns of Q past k are zero
        const int nab = (k + 15) / 16;
        for (int pass = 0; pass < 2 && k > 0; ++pass) {
            v4d w[RB];
#pragma unroll
            for (int ab = 0; ab < RB; ++ab) w[ab] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int ab = 0; ab < RB; ++ab) {
                if (ab >= nab) break;
#pragma unroll
                for (int ks = 0; ks < KS; ++ks)
                    w[ab] = __builtin_amdgcn_mfma_f64_16x16x4f64(L.Q[(16 * ab + lr) * LQ + 4 * ks + lg],
                                                                 y[ks >> 2][ks & 3], w[ab], 0, 0, 0);
            }
#pragma unroll
            for (int ab = 0; ab < RB; ++ab) {
                if (ab >= nab) break;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int ks = 4 * ab + e;
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb)
                        y[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(L.Q[(4 * ks + lg) * LQ + 16 * rb + lr],
                                                                     -w[ab][e], y[rb], 0, 0, 0);
                }
            }
        }
        // the block's columns to LDS, one row per lane from here on
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int e = 0; e < 4; ++e) L.S[lr * LQ + 16 * rb + 4 * e + lg] = y[rb][e];
        __syncthreads();
        // column-by-column decisions on explicitly projected vectors; lane q owns row q
        double qb[kBS];
        int na = 0;
        const int capleft = maxRank - k;
        for (int j = 0; j < kBS; ++j) {
            double yv = own ? L.S[j * LQ + tid] : 0.0;
#pragma unroll
            for (int pass = 0; pass < 2; ++pass)
#pragma unroll
                for (int a = 0; a < kBS; ++a)
                    if (a < na) {
                        const double d = wave_sum_dpp(qb[a] * yv);
                        yv -= qb[a] * d;
                    }
            const double nu = sqrt(wave_sum_dpp(yv * yv));
            if (nu > tau) {
                if (na >= capleft) { capped = 1; break; }
                const double qn = yv / nu;
#pragma unroll
                for (int a = 0; a < kBS; ++a)
                    if (a == na) qb[a] = qn;
                ++na;
                small_run = 0;
            } else if (++small_run >= r) {
                converged = 1;
                break;
            }
        }
        if (na > 0) {
            if (k > 0) {  // accepted columns -= Q (Q^T accepted), on the matrix cores
                __syncthreads();
#pragma unroll
                for (int a = 0; a < kBS; ++a)
                    if (own) L.S[a * LQ + tid] = a < na ? qb[a] : 0.0;
                __syncthreads();
                v4d w[RB], qa[RB];
#pragma unroll
                for (int ab = 0; ab < RB; ++ab) w[ab] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int ab = 0; ab < RB; ++ab) {
                    if (ab >= nab) break;
#pragma unroll
                    for (int ks = 0; ks < KS; ++ks)
                        w[ab] = __builtin_amdgcn_mfma_f64_16x16x4f64(L.Q[(16 * ab + lr) * LQ + 4 * ks + lg],
                                                                     L.S[lr * LQ + 4 * ks + lg], w[ab], 0, 0, 0);
                }
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                    for (int e = 0; e < 4; ++e) qa[rb][e] = L.S[lr * LQ + 16 * rb + 4 * e + lg];
#pragma unroll
                for (int ab = 0; ab < RB; ++ab) {
                    if (ab >= nab) break;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int ks = 4 * ab + e;
#pragma unroll
                        for (int rb = 0; rb < RB; ++rb)
                            qa[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(L.Q[(4 * ks + lg) * LQ + 16 * rb + lr],
                                                                          -w[ab][e], qa[rb], 0, 0, 0);
                    }
                }
                __syncthreads();
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                    for (int e = 0; e < 4; ++e) L.S[lr * LQ + 16 * rb + 4 * e + lg] = qa[rb][e];
                __syncthreads();
#pragma unroll
                for (int a = 0; a < kBS; ++a)
                    if (a < na) qb[a] = own ? L.S[a * LQ + tid] : 0.0;
            }
            // re-normalise within the block (MGS) and append to Q
#pragma unroll
            for (int j = 0; j < kBS; ++j)
                if (j < na) {
                    double q = qb[j];
#pragma unroll
                    for (int a = 0; a < j; ++a) {
                        const double d = wave_sum_dpp(qb[a] * q);
                        q -= qb[a] * d;
                    }
                    const double nu = sqrt(wave_sum_dpp(q * q));
                    q /= nu;
                    qb[j] = q;
                    if (own) L.Q[(k + j) * LQ + tid] = q;
                }
        }
        k += na;
        __syncthreads();
    }
    // U = Q, V = A^T Q: V[c][a] = sum_q A[q][c] Q[q][a]; lane l holds
    // A^T[16 cb + lr][4 ks + lg] = A[4 ks + lg][16 cb + lr] and Q[4 ks + lg][16 ab + lr]
    double* Us = Uslot + slot * (size_t)B * slotCols;
    double* Vs = Vslot + slot * (size_t)B * slotCols;
    for (int e = tid; e < k * B; e += 64) Us[e] = L.Q[(e / B) * LQ + e % B];
    const int nab = (k + 15) / 16;
    for (int cb = 0; cb < RB; ++cb) {
        const int c = 16 * cb + lr;
        v4d acc[RB];
#pragma unroll
        for (int ab = 0; ab < RB; ++ab) acc[ab] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
        for (int ks = 0; ks < KS; ++ks) {
            const int q = 4 * ks + lg;
            const double a = entry(&L.xr[q * 3], q, c);
#pragma unroll
            for (int ab = 0; ab < RB; ++ab) {
                if (ab >= nab) break;
                acc[ab] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, L.Q[(16 * ab + lr) * LQ + q], acc[ab], 0, 0, 0);
            }
        }
#pragma unroll
        for (int ab = 0; ab < RB; ++ab) {
            if (ab >= nab) break;
            const int a = 16 * ab + lr;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int cc = 16 * cb + 4 * e + lg;
                if (a < k && cc < B) Vs[(size_t)a * B + cc] = acc[ab][e];
            }
        }
    }
    if (tid == 0) {
        ranks[slot] = k;
        caps[slot] = (capped || (!converged && k >= maxRank)) ? 1 : 0;
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

template <int B, int D>
static void launch_tlr(Context& c, const TlrArgs& g, uint64_t t0, uint64_t nt, double* Us, double* Vs, int slotCols,
                       int32_t* ranks, int32_t* caps, double* diag) {
    const size_t lds = sizeof(TileLds<B>);
    auto kern = k_tlr_ara<B, D>;
    if (lds > 64 * 1024) HMX_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    kern<<<(unsigned)nt, 64, lds, c.stream>>>(g, t0, Us, Vs, slotCols, ranks, caps, diag);
    HMX_LAUNCHED();
}

template <int B>
static void launch_tlr_d(Context& c, uint32_t d, const TlrArgs& g, uint64_t t0, uint64_t nt, double* Us, double* Vs,
                         int slotCols, int32_t* ranks, int32_t* caps, double* diag) {
    if (d == 1) launch_tlr<B, 1>(c, g, t0, nt, Us, Vs, slotCols, ranks, caps, diag);
    else if (d == 2) launch_tlr<B, 2>(c, g, t0, nt, Us, Vs, slotCols, ranks, caps, diag);
    else launch_tlr<B, 3>(c, g, t0, nt, Us, Vs, slotCols, ranks, caps, diag);
}

static void launch_tlr_b(Context& c, uint32_t b, uint32_t d, const TlrArgs& g, uint64_t t0, uint64_t nt, double* Us,
                         double* Vs, int slotCols, int32_t* ranks, int32_t* caps, double* diag) {
    switch (b) {
        case 8: launch_tlr_d<8>(c, d, g, t0, nt, Us, Vs, slotCols, ranks, caps, diag); break;
        case 16: launch_tlr_d<16>(c, d, g, t0, nt, Us, Vs, slotCols, ranks, caps, diag); break;
        case 32: launch_tlr_d<32>(c, d, g, t0, nt, Us, Vs, slotCols, ranks, caps, diag); break;
        default: launch_tlr_d<64>(c, d, g, t0, nt, Us, Vs, slotCols, ranks, caps, diag); break;
    }
}

double ara_threshold(double tol) { return tol / (10.0 * std::sqrt(2.0 / 3.141592653589793)); }

// Compress 'total' tiles in batches sized so the per-tile slots stay under
// ~2 GiB; ranks[total] receives the ranks, U/V the packed factors in tile
// order.  slotCols: the largest rank cap of the launch's clouds.  Returns
// the number of capped tiles (and each tile's flag in capsOut when given).
static uint64_t compress_tiles(Context& c, uint32_t b, uint32_t d, const TlrArgs& g, int slotCols, uint64_t total,
                               int32_t* ranks, double* diag, DBuf<double>& U, DBuf<double>& V, uint64_t& rankSum,
                               std::vector<int32_t>* capsOut = nullptr) {
    if (capsOut) capsOut->assign(total, 0);
    const size_t slotBytes = 2 * (size_t)b * slotCols * sizeof(double);
    uint64_t batch = std::max<uint64_t>(1, ((size_t)2 << 30) / std::max<size_t>(slotBytes, 1));
    batch = std::min<uint64_t>(batch, std::max<uint64_t>(total, 1));
    DBuf<double> Us(batch * b * slotCols), Vs(batch * b * slotCols);
    DBuf<int32_t> caps(batch);
    DBuf<int64_t> boff(batch + 1);
    std::vector<DBuf<double>> partsU, partsV;
    std::vector<uint64_t> partBase;
    std::vector<int32_t> hcaps(batch);
    uint64_t running = 0, capHits = 0;
    for (uint64_t t0 = 0; t0 < total; t0 += batch) {
        const uint64_t nt = std::min<uint64_t>(batch, total - t0);
        int32_t* rk = ranks + t0;
        launch_tlr_b(c, b, d, g, t0, nt, Us.p, Vs.p, slotCols, rk, caps.p, diag);
        exclusive_scan_i32_i64(c, rk, boff.p, nt);
        int64_t btotal = 0;
        HMX_HIP(hipMemcpyAsync(&btotal, boff.p + nt, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
        HMX_HIP(hipMemcpyAsync(hcaps.data(), caps.p, nt * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
        HMX_HIP(hipStreamSynchronize(c.stream));
        for (uint64_t i = 0; i < nt; ++i) capHits += (uint64_t)hcaps[i];
        if (capsOut) std::copy(hcaps.begin(), hcaps.begin() + nt, capsOut->begin() + t0);
        partsU.emplace_back((size_t)btotal * b);
        partsV.emplace_back((size_t)btotal * b);
        partBase.push_back(running);
        if (btotal > 0) {
            k_compact<<<(unsigned)((nt + 3) / 4), 256, 0, c.stream>>>(nt, (int)b, slotCols, rk, boff.p, Us.p, Vs.p,
                                                                      partsU.back().p, partsV.back().p);
            HMX_LAUNCHED();
        }
        running += (uint64_t)btotal;
        HMX_HIP(hipStreamSynchronize(c.stream));
    }
    Us.release();
    Vs.release();
    rankSum = running;
    if (partsU.size() == 1) {  // one launch: its packed factors are the result
        U = std::move(partsU[0]);
        V = std::move(partsV[0]);
        return capHits;
    }
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

static TlrCloud tlr_cloud(const hmx_config& cfg, uint32_t b, uint32_t d, uint32_t S, int64_t leafBase) {
    TlrCloud q;
    q.leafBase = leafBase;
    q.first = 0;
    q.S = S;
    q.maxRank = cfg.maxRankPolicy == HMX_MAXRANK_REF ? (int)b / 2 : (int)b;
    q.ell = cfg.correlationLength > 0 ? cfg.correlationLength : default_ell(d);
    q.tau = ara_threshold((double)cfg.lowestLevelTolerance);
    return q;
}

static void check_tlr_args(const hmx_config& cfg, uint32_t b, uint32_t S) {
    HMX_CHECK(b == 8 || b == 16 || b == 32 || b == 64, HMX_EUNSUPPORTED,
              "TLR tiles: bucketSize must be 8, 16, 32 or 64");
    HMX_CHECK(cfg.araBlockSize == kBS, HMX_EUNSUPPORTED, "ARA block size must be 16");
    HMX_CHECK(cfg.araR >= 1, HMX_EINVAL, "araR must be >= 1");
    HMX_CHECK(S >= 2, HMX_EINVAL, "TLR needs at least two segments (n > bucketSize)");
}

void build_tlr_cm(Context& c, const hmx_config& cfg, const KDTree& t, const double* d_points, TLR& out) {
    check_tlr_args(cfg, t.b, t.numSegments);
    const uint32_t b = t.b, S = t.numSegments;
    out.ordering = HMX_COLUMN_MAJOR;
    out.S = S;
    out.b = b;
    out.n = t.n;
    const uint64_t total = (uint64_t)S * S;
    out.ranks.alloc(total);
    out.offsets.alloc(total + 1);
    out.diagonal.alloc((size_t)S * b * b);
    const TlrCloud q = tlr_cloud(cfg, b, t.d, S, 0);
    DBuf<TlrCloud> dq(1);
    HMX_HIP(hipMemcpyAsync(dq.p, &q, sizeof(TlrCloud), hipMemcpyHostToDevice, c.stream));
    TlrArgs g{};
    g.n = t.n;
    g.p = d_points;
    g.perm = t.perm.p;
    g.offsets = t.offsets.p;
    g.clouds = dq.p;
    g.r = cfg.araR;
    g.seed = cfg.seed;
    uint64_t rankSum = 0;
    out.capHits = compress_tiles(c, b, t.d, g, q.maxRank, total, out.ranks.p, out.diagonal.p, out.U, out.V, rankSum);
    out.rankSum = rankSum;
    exclusive_scan_i32_i64(c, out.ranks.p, out.offsets.p, total);
    HMX_HIP(hipStreamSynchronize(c.stream));
}

// One piece of a p x p partition of the TLR matrix (the reference's
// large-matrices idea, README.md:4: the tiles of rows [i S/p, (i+1) S/p) and
// columns [j S/p, (j+1) S/p), so a matrix beyond memory can be built piece
// by piece).  Tiles in column-major order within the piece; every tile is
// the full matrix's tile bit for bit (same kernel, same random stream
// u*S + v); diagonal pieces carry their diagonal blocks.
void build_tlr_piece(Context& c, const hmx_config& cfg, const KDTree& t, const double* d_points, uint32_t pieces,
                     uint32_t rowPiece, uint32_t colPiece, TLR& out) {
    check_tlr_args(cfg, t.b, t.numSegments);
    const uint32_t b = t.b, S = t.numSegments;
    HMX_CHECK(pieces >= 1 && S % pieces == 0, HMX_EINVAL, "TLR pieces: the leaf count must divide into pieces");
    HMX_CHECK(rowPiece < pieces && colPiece < pieces, HMX_EINVAL, "TLR pieces: piece index out of range");
    const uint32_t Sp = S / pieces;
    out.ordering = HMX_COLUMN_MAJOR;
    out.S = Sp;
    out.b = b;
    out.n = t.n;
    out.pieces = pieces;
    out.rowPiece = rowPiece;
    out.colPiece = colPiece;
    const uint64_t total = (uint64_t)Sp * Sp;
    std::vector<int32_t> codes(total);
    for (uint32_t uu = 0; uu < Sp; ++uu)
        for (uint32_t vv = 0; vv < Sp; ++vv)
            codes[(size_t)uu * Sp + vv] = (int32_t)morton_xy(colPiece * Sp + uu, rowPiece * Sp + vv);
    out.ranks.alloc(total);
    out.offsets.alloc(total + 1);
    const bool onDiag = rowPiece == colPiece;
    if (onDiag) out.diagonal.alloc((size_t)Sp * b * b);
    const TlrCloud q = tlr_cloud(cfg, b, t.d, S, 0);
    DBuf<TlrCloud> dq(1);
    DBuf<int32_t> d_codes(total);
    HMX_HIP(hipMemcpyAsync(dq.p, &q, sizeof(TlrCloud), hipMemcpyHostToDevice, c.stream));
    HMX_HIP(hipMemcpyAsync(d_codes.p, codes.data(), total * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
    TlrArgs g{};
    g.n = t.n;
    g.p = d_points;
    g.perm = t.perm.p;
    g.offsets = t.offsets.p;
    g.clouds = dq.p;
    g.tileMO = d_codes.p;
    g.diagLeaf0 = colPiece * Sp;
    g.r = cfg.araR;
    g.seed = cfg.seed;
    uint64_t rankSum = 0;
    out.capHits = compress_tiles(c, b, t.d, g, q.maxRank, total, out.ranks.p, onDiag ? out.diagonal.p : nullptr, out.U,
                                 out.V, rankSum);
    out.rankSum = rankSum;
    exclusive_scan_i32_i64(c, out.ranks.p, out.offsets.p, total);
    HMX_HIP(hipStreamSynchronize(c.stream));
}

// Leaf-level tiles of one or more clouds compressed straight from the kernel
// (the TLR ARA over each cloud's Morton tile list, points read from the
// padded leaf space P): the same factors, bit for bit, as the TLR matrix's
// tiles at those positions.  One launch sequence for all jobs; each job's
// level gets its range of the packed factors as a view.
// leaf tiles of a symmetric-tiles build laid out in the level: output tile j
// takes compressed tile src's k columns (b rows each), U and V exchanged when
// swap is set (the transposed partner); one grid row per output tile
struct LeafPlace {
    int64_t dst, src;
    int32_t k, swap;
};
__global__ void __launch_bounds__(256) k_leaf_place(int b, const LeafPlace* __restrict__ pl,
                                                    const double* __restrict__ Uc, const double* __restrict__ Vc,
                                                    double* __restrict__ U, double* __restrict__ V) {
    const LeafPlace q = pl[blockIdx.y];
    const size_t cnt = (size_t)q.k * b, d = (size_t)q.dst * b, s = (size_t)q.src * b;
    const double* su = q.swap ? Vc : Uc;
    const double* sv = q.swap ? Uc : Vc;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < cnt; e += (size_t)gridDim.x * 256) {
        U[d + e] = su[s + e];
        V[d + e] = sv[s + e];
    }
}

void compress_leaf_tiles_padded(Context& c, uint32_t b, uint32_t d, const double* P,
                                const std::vector<LeafJob>& jobs) {
    const int nj = (int)jobs.size();
    if (nj == 0) return;
    std::vector<TlrCloud> cl(nj);
    std::vector<int32_t> tiles, tileCloud;
    // symmetric-tiles jobs compress only their tiles (u, v) with u < v (the
    // others are the transposes); outSrc maps every output tile of every job
    // to its compressed tile (~i: exchanged)
    std::vector<int32_t> outSrc;
    bool anyMirror = false;
    int slotCols = 1;
    for (int j = 0; j < nj; ++j) {
        const LeafJob& J = jobs[j];
        check_tlr_args(*J.cfg, b, J.S);
        cl[j] = tlr_cloud(*J.cfg, b, d, J.S, J.leafBase);
        slotCols = std::max(slotCols, cl[j].maxRank);
        const int32_t cb = (int32_t)tiles.size();
        if (J.cfg->symmetricTiles) {
            std::vector<int> prim;
            std::vector<int32_t> src1;
            mirror_plan(*J.tiles, prim, src1);
            for (int t : prim) tiles.push_back((*J.tiles)[t]);
            tileCloud.insert(tileCloud.end(), prim.size(), j);
            for (int32_t q : src1) outSrc.push_back(q >= 0 ? q + cb : ~(~q + cb));
            anyMirror = true;
        } else {
            tiles.insert(tiles.end(), J.tiles->begin(), J.tiles->end());
            tileCloud.insert(tileCloud.end(), J.tiles->size(), j);
            for (int32_t q = 0; q < (int32_t)J.tiles->size(); ++q) outSrc.push_back(cb + q);
        }
        HLevel& out = *J.out;
        out.numTiles = (int)J.tiles->size();
        out.m = (int)b;
        out.maxRank = cl[j].maxRank;
        out.tol = J.cfg->lowestLevelTolerance;
        out.h_ranks.assign(out.numTiles, 0);
        out.h_offsets.assign(out.numTiles + 1, 0);
        out.rankSum = 0;
        out.capHits = 0;
        out.U.release();
        out.V.release();
    }
    const uint64_t T = tiles.size();
    if (T == 0) return;
    DBuf<TlrCloud> dq(nj);
    DBuf<int32_t> d_tiles(T), d_tc(T), d_r(T);
    HMX_HIP(hipMemcpyAsync(dq.p, cl.data(), nj * sizeof(TlrCloud), hipMemcpyHostToDevice, c.stream));
    HMX_HIP(hipMemcpyAsync(d_tiles.p, tiles.data(), T * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
    HMX_HIP(hipMemcpyAsync(d_tc.p, tileCloud.data(), T * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
    TlrArgs g{};
    g.P = reinterpret_cast<const double4*>(P);
    g.clouds = dq.p;
    g.tileCloud = nj > 1 ? d_tc.p : nullptr;
    g.tileMO = d_tiles.p;
    g.r = jobs[0].cfg->araR;
    g.seed = jobs[0].cfg->seed;
    auto U = std::make_shared<DBuf<double>>(), V = std::make_shared<DBuf<double>>();
    uint64_t rankSum = 0;
    std::vector<int32_t> hr(T), hcap;
    compress_tiles(c, b, d, g, slotCols, T, d_r.p, nullptr, *U, *V, rankSum, &hcap);
    HMX_HIP(hipMemcpyAsync(hr.data(), d_r.p, T * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    HMX_HIP(hipStreamSynchronize(c.stream));
    if (anyMirror) {  // lay out every job's level from the compressed tiles
        std::vector<int64_t> coff(T + 1, 0);
        for (uint64_t i = 0; i < T; ++i) coff[i + 1] = coff[i] + hr[i];
        std::vector<LeafPlace> place;
        int64_t acc = 0;
        size_t o = 0;
        for (int j = 0; j < nj; ++j) {
            HLevel& out = *jobs[j].out;
            const int64_t base = acc;
            for (int q = 0; q < out.numTiles; ++q, ++o) {
                const int32_t s = outSrc[o];
                const int i = s >= 0 ? s : ~s;
                const int32_t rk = hr[i];
                out.h_ranks[q] = rk;
                out.h_offsets[q] = acc - base;
                out.capHits += hcap[i];
                if (rk > 0) place.push_back(LeafPlace{acc, coff[i], rk, s < 0 ? 1 : 0});
                acc += rk;
            }
            out.h_offsets[out.numTiles] = acc - base;
            out.rankSum = acc - base;
        }
        auto U2 = std::make_shared<DBuf<double>>((size_t)std::max<int64_t>(1, acc) * b);
        auto V2 = std::make_shared<DBuf<double>>((size_t)std::max<int64_t>(1, acc) * b);
        if (!place.empty()) {
            DBuf<LeafPlace> dp(place.size());
            HMX_HIP(hipMemcpyAsync(dp.p, place.data(), place.size() * sizeof(LeafPlace), hipMemcpyHostToDevice,
                                   c.stream));
            for (size_t p0 = 0; p0 < place.size(); p0 += 65535) {  // grid rows per launch
                const unsigned rows = (unsigned)std::min<size_t>(65535, place.size() - p0);
                k_leaf_place<<<dim3(std::max(1u, (unsigned)((slotCols * b + 255) / 256)), rows), 256, 0, c.stream>>>(
                    (int)b, dp.p + p0, U->p, V->p, U2->p, V2->p);
                HMX_LAUNCHED();
            }
            HMX_HIP(hipStreamSynchronize(c.stream));
        }
        int64_t base = 0;
        for (int j = 0; j < nj; ++j) {
            HLevel& out = *jobs[j].out;
            out.U.view(U2, (size_t)base * b, (size_t)out.rankSum * b);
            out.V.view(V2, (size_t)base * b, (size_t)out.rankSum * b);
            base += out.rankSum;
        }
        return;
    }
    int64_t acc = 0;
    size_t t0 = 0;
    for (int j = 0; j < nj; ++j) {
        HLevel& out = *jobs[j].out;
        const int64_t base = acc;
        for (int q = 0; q < out.numTiles; ++q) {
            const int32_t rk = hr[t0 + q];
            out.h_ranks[q] = rk;
            out.h_offsets[q] = acc - base;
            acc += rk;
            out.capHits += hcap[t0 + q];
        }
        out.h_offsets[out.numTiles] = acc - base;
        out.rankSum = acc - base;
        out.U.view(U, (size_t)base * b, (size_t)out.rankSum * b);
        out.V.view(V, (size_t)base * b, (size_t)out.rankSum * b);
        t0 += out.numTiles;
    }
    HMX_CHECK((uint64_t)acc == rankSum, HMX_EDEVICE, "leaf tile compression: rank sum mismatch");
}

// Dense diagonal blocks, exactly as the TLR kernel writes them (D[leaf] is
// column-major b x b, K(x_c, x_q) for real points, identity on the padding),
// from the padded leaf space of one or more clouds: identical coordinates
// and operations, so identical bits.
template <int D>
__global__ void __launch_bounds__(256) k_dense_diag_padded(int64_t total, int b, const double4* __restrict__ P,
                                                           const TlrCloud* __restrict__ clouds, int nclouds,
                                                           double* __restrict__ diag) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= total) return;
    const int64_t bb = (int64_t)b * b;
    const int64_t k = e / bb;  // the launch's k-th leaf
    const int c = (int)((e % bb) / b), q = (int)(e % b);
    int lo = 0, hi = nclouds - 1;  // the last cloud starting at or before it
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (clouds[mid].first <= k) lo = mid;
        else hi = mid - 1;
    }
    const int64_t leaf = clouds[lo].leafBase + (k - clouds[lo].first);
    const double4 pc = P[leaf * b + c], pq = P[leaf * b + q];
    double val;
    if (pc.w != 0.0 && pq.w != 0.0) {
        const double xc[3] = {pc.x, pc.y, pc.z}, xq[3] = {pq.x, pq.y, pq.z};
        val = kernel_eval<D>(xc, xq, clouds[lo].ell);
    } else {
        val = c == q ? 1.0 : 0.0;
    }
    diag[leaf * bb + (e % bb)] = val;
}

void build_dense_diagonal_padded(Context& c, uint32_t b, uint32_t d, const double* P, const std::vector<LeafJob>& jobs,
                                 double* diag) {
    const int nj = (int)jobs.size();
    if (nj == 0) return;
    std::vector<TlrCloud> cl(nj);
    int64_t leaves = 0;
    for (int j = 0; j < nj; ++j) {
        cl[j] = tlr_cloud(*jobs[j].cfg, b, d, jobs[j].S, jobs[j].leafBase);
        cl[j].first = leaves;
        leaves += jobs[j].S;
    }
    DBuf<TlrCloud> dq(nj);
    HMX_HIP(hipMemcpyAsync(dq.p, cl.data(), nj * sizeof(TlrCloud), hipMemcpyHostToDevice, c.stream));
    const int64_t tot = leaves * b * b;
    const unsigned g = (unsigned)((tot + 255) / 256);
    const double4* P4 = reinterpret_cast<const double4*>(P);
    if (d == 1) k_dense_diag_padded<1><<<g, 256, 0, c.stream>>>(tot, (int)b, P4, dq.p, nj, diag);
    else if (d == 2) k_dense_diag_padded<2><<<g, 256, 0, c.stream>>>(tot, (int)b, P4, dq.p, nj, diag);
    else k_dense_diag_padded<3><<<g, 256, 0, c.stream>>>(tot, (int)b, P4, dq.p, nj, diag);
    HMX_LAUNCHED();
    HMX_HIP(hipStreamSynchronize(c.stream));  // the cloud table goes out of scope
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
